"""Batched passes (pivx_batch_*): many pairs of uneven frame size with
non-square windows run with one persistent work-queue correlation launch per
pass; every pair's raw candidates, filtered fields and median-only field must
be bit-identical to the same pair run alone with pivx_run_pair, eagerly, as a
replayed graph, and after the autotuner has picked the launch shape."""
import numpy as np
import pytest

from test_gpu_parity import _report

pytestmark = pytest.mark.gpu

SIZES = [(1024, 1024, 0.05), (2048, 1536, 0.02), (1280, 720, 0.08), (1024, 1024, 0.03), (1536, 1024, 0.07)]


def _params(pivx, nx, ny):
    return pivx.init_params(nx=nx, ny=ny, ws_x1=64, ws_y1=32, spacing_x1=32, spacing_y1=16, npass1=2, ws_x2=32,
                            ws_y2=16, spacing_x2=16, spacing_y2=8, npass2=2)


def _fields(x):
    return [x.raw(1), x.peaks(1), x.filtered(1), x.raw(2), x.peaks(2), x.filtered(2), x.filtered(3)]


def _equal(A, B):
    for u, w in zip(A, B):
        for a, b in zip(u, w):
            np.testing.assert_array_equal(a, b)


def test_batch_equals_single_pairs(pivx):
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    xs = []
    try:
        for i, (nx, ny, dens) in enumerate(SIZES):
            x = pivx.Xcorr2dMulti(_params(pivx, nx, ny))
            xs.append(x)
            x.synth_pair_device(pivx.synth_params(nx, ny, seed=4000 + i, density=dens, flow=2, a0=10.0))
        alone = []
        for x in xs:
            x.run_pair(fp, fp)
            alone.append(_fields(x))
        with pivx.Batch(xs) as b:
            for graph in (False, True):
                b.set_graph_mode(graph)
                for _ in range(2):
                    b.run(fp, fp)
                    b.synchronize()
                    for x, want in zip(xs, alone):
                        _equal(_fields(x), want)
            b.autotune(2)
            t1, t2 = b.tuning(1), b.tuning(2)
            b.run(fp, fp)
            b.synchronize()
            for x, want in zip(xs, alone):
                _equal(_fields(x), want)
            _report("batch", pairs=len(xs), grid1=t1, grid2=t2)
            assert t1["items"] > 0 and t2["items"] > 0
    finally:
        for x in xs:
            x.close()


# config-4 frame sizes with 64x64 / 32x32 windows (the two-wave 64x64
# correlation kernel and its persistent form)
SIZES_64 = [(4096, 2048, 0.05), (1024, 1024, 0.02), (6600, 4400, 0.05), (2048, 1536, 0.08)]


def test_batch_equals_single_pairs_64x64_large_frames(pivx):
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    xs = []
    try:
        for i, (nx, ny, dens) in enumerate(SIZES_64):
            p = pivx.init_params(nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2,
                                 ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2)
            x = pivx.Xcorr2dMulti(p)
            xs.append(x)
            x.synth_pair_device(pivx.synth_params(nx, ny, seed=5000 + i, density=dens, flow=2, a0=10.0))
        alone = []
        for x in xs:
            x.run_pair(fp, fp)
            alone.append(_fields(x))
        with pivx.Batch(xs) as b:
            b.autotune(1)
            for graph in (False, True):
                b.set_graph_mode(graph)
                b.run(fp, fp)
                b.synchronize()
                for x, want in zip(xs, alone):
                    _equal(_fields(x), want)
            _report("batch 64x64", pairs=len(xs), grid1=b.tuning(1), grid2=b.tuning(2))
    finally:
        for x in xs:
            x.close()


def test_batch_rejects_mixed_windows(pivx):
    a = pivx.Xcorr2dMulti(_params(pivx, 1024, 1024))
    p = _params(pivx, 1024, 1024)
    p.ws_x2, p.ws_y2 = 32, 32
    b = pivx.Xcorr2dMulti(p)
    try:
        with pytest.raises(pivx.PivError, match="window sizes"):
            pivx.Batch([a, b])
        q = pivx.init_params(nx=1024, ny=1024, ws_x1=16, ws_y1=16, spacing_x1=8, spacing_y1=8)
        with pivx.Xcorr2dMulti(q) as c:
            with pytest.raises(pivx.PivError, match="no batched correlation kernel"):
                pivx.Batch([c])
    finally:
        a.close()
        b.close()
