import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def rel_err(a, b):
    """Per-component error normalised by max(|b|, max over cells |b_n|):
    |a-b| / max(|b|, max_cells |b_n|) (component axis first)."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    out = 0.0
    for n in range(a.shape[0]):
        scale = np.maximum(np.abs(b[n]), np.max(np.abs(b[n])))
        scale = np.where(scale == 0.0, 1.0, scale)
        out = max(out, float(np.max(np.abs(a[n] - b[n]) / scale)))
    return out
