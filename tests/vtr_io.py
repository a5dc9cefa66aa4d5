"""Reading an ASCII VTK RectilinearGrid file into its structure lines and the
float32 values of each DataArray (the comparison unit of the VTR tests)."""

import numpy as np


def parse(path):
    """Returns (tags, arrays): the non-numeric lines stripped of trailing
    blanks, and {array name: float32 values}."""
    tags, arrays = [], {}
    name, vals = None, []
    with open(path) as fh:
        for line in fh:
            t = line.rstrip()
            if t.lstrip().startswith("<"):
                if name is not None and t.strip() == "</DataArray>":
                    text = " ".join(vals).split()
                    arrays[name] = np.array([float(v) for v in text], dtype=np.float64).astype(
                        np.float32) if len(text) < 100000 else \
                        np.array(text, dtype=np.float64).astype(np.float32)
                    name, vals = None, []
                elif 'Name="' in t:
                    name = t.split('Name="')[1].split('"')[0]
                tags.append(t)
            elif t.strip():
                vals.append(t)
    return tags, arrays
