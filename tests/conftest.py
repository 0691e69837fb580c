"""Shared test helpers: load the engine binding and the oracle, build the
seeded synthetic inputs of the measured configurations."""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "piv_xcorr2d_v1.0_amd")
ORACLE_DIR = os.path.join(REPO, "oracle")
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_pivx():
    if "pivx" in sys.modules:
        return sys.modules["pivx"]
    spec = importlib.util.spec_from_file_location("pivx", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pivx"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    if ORACLE_DIR not in sys.path:
        sys.path.insert(0, ORACLE_DIR)
    import pyoracle  # noqa: E402

    return pyoracle


@pytest.fixture(scope="session")
def pivx():
    return load_pivx()


@pytest.fixture(scope="session")
def orc():
    return load_oracle()


def tie_margin(ws_x, ws_y):
    """Relative margin under which two correlation values count as tied: a
    few FP32 roundings per FFT stage, 8 * 2^-23 * log2(N)."""
    return 8.0 * 2.0 ** -23 * np.log2(ws_x * ws_y)


def config1_params(pivx, **kw):
    """512x512, 32x32 windows at spacing 16, one pass, top-4, max_disp 0.5."""
    d = dict(nx=512, ny=512, ws_x1=32, ws_y1=32, spacing_x1=16, spacing_y1=16, npass1=1,
             ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=0, npeak_save=4,
             max_dispx_percent1=0.5, max_dispy_percent1=0.5, max_dispx_percent2=0.5, max_dispy_percent2=0.5)
    d.update(kw)
    return pivx.init_params(**d)


def config1_pair(pivx, seed=1, nx=512, ny=512, u=3.3, v=-1.7):
    sp = pivx.synth_params(nx, ny, seed=seed, flow=0, a0=u, a1=v)
    return pivx.synth_pair(sp)


def geom_dict(g):
    return dict(nx_vec=g.nx_vec, ny_vec=g.ny_vec, border_x=g.border_x, border_y=g.border_y,
                npeak_slots=g.npeak_slots, left_x=g.left_x, left_y=g.left_y, right_x=g.right_x,
                right_y=g.right_y, nx_aug=g.nx_aug, ny_aug=g.ny_aug)
