"""The C ABI library loads and exports every function include/simco_proj.h
declares; argument validation and status strings work without a GPU."""

import ctypes
import os
import re

import pytest

from simcoflow3d_amd import engine
from simcoflow3d_amd._paths import ENGINE_LIB, INCLUDE_DIR


def declared_functions():
    text = open(os.path.join(INCLUDE_DIR, "simco_proj.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_binding_symbols():
    decl = declared_functions()
    assert set(engine.EXPORTED_SYMBOLS) <= set(decl)
    assert len(decl) >= 15


def test_library_exports_every_declared_function():
    assert os.path.exists(ENGINE_LIB), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(ENGINE_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_defaults():
    L = engine.load()
    for s in range(0, -10, -1):
        assert L.sp_strerror(s) != b"unknown status"
        assert L.sp_strerror(s)
    p = engine.SpParams()
    L.sp_params_default(ctypes.byref(p))
    assert p.mg_pre == 2 and p.mg_post == 2 and p.div_tol == 1e-10 and p.mg_omega == 1.25
    assert p.mg_omega_coarse == 1.5
    assert p.rho_l == 998.3 and p.rho_g == 1.205 and p.epsi == 1e-3


def test_invalid_arguments_are_rejected_without_a_device():
    L = engine.load()
    h = ctypes.c_void_p()
    assert L.sp_create(None, None, 0) == engine.SP_ERR_INVALID_ARG
    p = engine.SpParams()
    L.sp_params_default(ctypes.byref(p))
    p.imax, p.jmax, p.kmax = 1, 8, 8  # grids need at least 2 cells per direction
    assert L.sp_create(ctypes.byref(h), ctypes.byref(p), 0) == engine.SP_ERR_INVALID_ARG
    p.imax, p.jmax, p.kmax = 1024, 1024, 512  # 2^29 cells: beyond 32-bit byte offsets of FP64 levels
    assert L.sp_create(ctypes.byref(h), ctypes.byref(p), 0) == engine.SP_ERR_INVALID_ARG
    p.imax, p.jmax, p.kmax = 8, 8, 8
    p.mg_omega = 2.0  # over-relaxation must stay in (0, 2)
    assert L.sp_create(ctypes.byref(h), ctypes.byref(p), 0) == engine.SP_ERR_INVALID_ARG
    p.mg_omega = 1.25
    p.precision = 7
    assert L.sp_create(ctypes.byref(h), ctypes.byref(p), 0) == engine.SP_ERR_INVALID_ARG
    assert L.sp_project(None, None, None, None) == engine.SP_ERR_INVALID_ARG
    assert L.sp_adjoint(None, None, None, None) == engine.SP_ERR_INVALID_ARG
