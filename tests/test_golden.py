"""The generator and the oracle reproduce the committed golden fixtures."""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs
from tests import golden_io


@pytest.mark.parametrize("name", golden_io.NAMES)
def test_generator_reproduces_fixture_inputs(name):
    cfg, c, _ = golden_io.load(name)
    g = configs.generate(cfg, size=c.shape)
    for k, v in c.fields.items():
        assert np.array_equal(g.fields[k], v), k


@pytest.mark.parametrize("name", golden_io.NAMES)
def test_oracle_reproduces_fixture_outputs(name):
    _, c, exp = golden_io.load(name)
    ref = orc.project(c, pin=(1, 1, 1))
    assert np.array_equal(ref.pocoef, exp["pocoef"])
    assert np.max(np.abs(ref.Pp - exp["p"])) <= 1e-12 * np.max(np.abs(exp["p"]))
    for k, a in (("u", ref.u), ("v", ref.v), ("w", ref.w)):
        assert np.max(np.abs(a - exp[k])) <= 1e-12 * max(1.0, np.max(np.abs(exp[k])))
