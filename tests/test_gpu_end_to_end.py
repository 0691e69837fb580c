"""GPU parity of the whole path and of the kernel paths the shape tests leave
out: odd vector-column counts (the idle second half-wave of the wave-pair
kernel, partial chunks of the four-window 16x16 kernel), the geometry the
engine derives at init, the hole counter of the fill stage, and the whole
two-grid pipeline against the FP64 oracle's own pipeline, at a mid size and
at the full 6600x4400 size.

End-to-end bar (stated per test): the GPU computes in FP32 and the oracle in
FP64, so the fields differ by rounding that each deformation pass carries
into the next one, and near-tied peaks or median statistics can pick another
candidate in a few windows.  The tests bound the RMS difference of the final
fields, the fraction of vectors farther apart than 0.01 px and the fraction
of validity flags that differ.
"""
import os

import numpy as np
import pytest

from conftest import config1_pair, config1_params, geom_dict, tie_margin
from test_gpu_parity import _report, check_pass, gpu_pass

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)

ODD_COLUMN_CASES = [
    # (nx, ny, ws, sp, nx_vec): the wave-pair kernel (windows up to 32x32, two per wave)
    (600, 400, (32, 32), (16, 16), 37),
    (588, 400, (24, 24), (12, 12), 49),
    (600, 400, (32, 16), (16, 8), 37),
    (600, 400, (16, 32), (8, 16), 75),
    # the four-window 16x16 kernel: nx_vec % 4 = 3, 2, 1
    (600, 320, (16, 16), (8, 8), 75),
    (592, 320, (16, 16), (8, 8), 74),
    (584, 320, (16, 16), (8, 8), 73),
]


@pytest.mark.parametrize("case", ODD_COLUMN_CASES)
def test_odd_vector_columns(pivx, orc, case):
    """Frames whose vector-column count leaves the last wave (or workgroup)
    of a row partly idle: undeformed and deformed passes against the FP64
    oracle, with the last column checked on its own (ncand, peak indices,
    validity, sub-pixel displacements)."""
    nx, ny, ws, sp, want_cols = case
    a, b = config1_pair(pivx, seed=41, nx=nx, ny=ny, u=2.2, v=-1.4)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=sp[0], spacing_y1=sp[1], npass1=2,
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4)
    rng = np.random.default_rng(43)
    with pivx.Xcorr2dMulti(p) as x:
        g = geom_dict(x.geometry(1))
        assert g["nx_vec"] == want_cols
        nt = g["nx_vec"] * g["ny_vec"]
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu0 = gpu_pass(x, 1)
        fx = (2.2 + 0.3 * rng.standard_normal(nt)).astype(np.float32)
        fy = (-1.4 + 0.3 * rng.standard_normal(nt)).astype(np.float32)
        x.set_filtered(1, fx, fy)
        x.extend_filtered_vecfield_p1()
        x.Calc_cc_map_p1(1)
        gpu1 = gpu_pass(x, 1)
        d = x.deformed(1)
    last = (np.arange(nt) % g["nx_vec"]) == g["nx_vec"] - 1
    for name, gpu, kw in (("pass 0", gpu0, dict(f2u8=b)),
                          ("deformed", gpu1, dict(f2def=d.astype(np.float64),
                                                  pred=(fx.astype(np.float64), fy.astype(np.float64))))):
        ref = orc.xcorr_pass(a, geom=g, ws=ws, sp=sp, K=4, tie_rel=tie_margin(*ws), threads=THREADS, **kw)
        check_pass(f"odd columns {nx}x{ny} {ws[0]}x{ws[1]} ({want_cols} cols) {name}", gpu, ref, ws, 4)
        sel = last & (ref["tie"] == 0)
        assert sel.sum() >= 0.9 * last.sum()
        np.testing.assert_array_equal(gpu["nc"][sel], ref["ncand"][sel])
        np.testing.assert_array_equal(gpu["idx"][sel], ref["pk_idx"][sel])
        np.testing.assert_array_equal(gpu["v"][sel], ref["valid"][sel])
        both = sel[:, None] & (gpu["v"] == 1)
        assert np.abs(gpu["dx"] - ref["dx"])[both].max() <= 1e-3
        assert np.abs(gpu["dy"] - ref["dy"])[both].max() <= 1e-3
        # the last column holds real vectors (the frame's right edge is covered)
        assert gpu["v"][last, 0].mean() > 0.5


GEOMETRY_CASES = [
    (512, 512, (32, 32), (16, 16), (32, 32), (16, 16)),     # config 1
    (6600, 4400, (64, 64), (64, 64), (32, 32), (16, 16)),   # config 2
    (600, 400, (32, 32), (16, 16), (16, 16), (8, 8)),       # odd column counts
    (2048, 2048, (64, 64), (32, 32), (32, 32), (8, 8)),     # 75% overlap on grid 2
    (301, 177, (20, 12), (50, 9), (24, 8), (12, 5)),        # spacing beyond the window
]


@pytest.mark.parametrize("case", GEOMETRY_CASES)
def test_device_geometry_equals_oracle(pivx, orc, case):
    """The geometry of a live handle (pads, aug sizes, borders, slot cap,
    sections) equals the oracle's and the host-only computation."""
    nx, ny, ws1, sp1, ws2, sp2 = case
    ip = pivx.init_params(nx=nx, ny=ny, ws_x1=ws1[0], ws_y1=ws1[1], spacing_x1=sp1[0], spacing_y1=sp1[1],
                          ws_x2=ws2[0], ws_y2=ws2[1], spacing_x2=sp2[0], spacing_y2=sp2[1])
    with pivx.Xcorr2dMulti(ip) as x:
        for grid, ws, sp in ((1, ws1, sp1), (2, ws2, sp2)):
            want = orc.geometry(nx, ny, 1, ws[0], ws[1], sp[0], sp[1], 0.4, 0.4)
            g = x.geometry(grid)
            h = pivx.compute_geometry(ip, grid)
            assert {k: getattr(g, k) for k in want} == want
            assert {k: getattr(h, k) for k in want} == want


def test_hole_counter_invariant(pivx):
    """The hole counter of the fill stage is 0 whenever the stream is idle and
    no hole is ever dropped, over every filter sequence: the pair sequence
    (eager and replayed graph), median-then-fill with one or three median
    passes, and fill without median passes (holes listed by the fill)."""
    nx, ny = 640, 480
    sp_ = pivx.synth_params(nx, ny, seed=8, flow=1, a0=100.0, a1=5.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2,
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                       max_dispy_percent2=0.4)
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    rng = np.random.default_rng(5)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)

        def clean(what):
            st = x.stats()
            _report("hole counter", after=what, pending=st["holes_pending"], drops=st["hole_list_drops"],
                    sweeps=st["fill_sweeps"])
            assert st["holes_pending"] == 0 and st["hole_list_drops"] == 0, what

        x.run_pair(fp, fp)
        clean("run_pair")
        x.set_graph_mode(True)
        x.run_pair(fp, fp)
        x.run_pair(fp, fp)
        clean("run_pair graph replays")
        nt = x.geometry(2).nt_vec
        dx = rng.normal(1.0, 0.2, (nt, 4)).astype(np.float32)
        dy = rng.normal(-1.0, 0.2, (nt, 4)).astype(np.float32)
        v = (rng.random((nt, 4)) > 0.3).astype(np.uint8)
        for npm in (1, 3, 0):
            x.set_raw(2, dx, dy, v)
            x.MedianFilter_Smoothing_p2(npm, 5, 2.0, 2.0, 1, 1)
            x.synchronize()
            clean(f"grid-2 filter with {npm} median passes")
            fv = x.filtered(2)[2]
            assert fv.all(), "fill leaves no hole in a field with valid vectors"
        x.MedianFilter_Smoothing_p1(1, 5, 2.0, 2.0, 0)
        clean("grid-1 filter, one median pass")


def _pair_fields(x):
    return dict(raw1=x.raw(1), filt1=x.filtered(1), raw2=x.raw(2), filt2=x.filtered(2), mo2=x.filtered(3))


def _compare_fields(name, gpu, ref, rms_tol, far_frac, valid_frac):
    """Final fields of the GPU pipeline against the oracle's pipeline: RMS of
    the displacement difference over vectors valid on both sides, the
    fraction farther apart than 0.01 px, and the fraction of differing
    validity flags."""
    out = {}
    for key in ("filt1", "filt2", "mo2"):
        gx, gy, gv = gpu[key]
        rx, ry, rv = ref[key]
        both = (gv == 1) & (rv == 1)
        d = np.hypot(gx[both] - rx[both], gy[both] - ry[both])
        rms = float(np.sqrt(np.mean(d ** 2))) if d.size else 0.0
        far = float(np.mean(d > 0.01)) if d.size else 0.0
        vdiff = float(np.mean(gv != rv))
        out[key] = (rms, far, vdiff)
        _report(name, field=key, vectors=len(gv), rms_diff_px=f"{rms:.2e}", frac_gt_0p01px=f"{far:.2e}",
                frac_valid_differs=f"{vdiff:.2e}", max_diff_px=f"{float(d.max()) if d.size else 0.0:.2e}")
    for key, (rms, far, vdiff) in out.items():
        assert rms <= rms_tol, (key, rms)
        assert far <= far_frac, (key, far)
        assert vdiff <= valid_frac, (key, vdiff)
    # raw candidate 0 of the last grid-2 pass
    gx, gy, gv = gpu["raw2"]
    rx, ry, rv = ref["raw2"]
    both = (gv[:, 0] == 1) & (rv[:, 0] == 1)
    d0 = np.hypot(gx[both, 0] - rx[both, 0], gy[both, 0] - ry[both, 0])
    _report(name, field="raw2[k=0]", rms_diff_px=f"{float(np.sqrt(np.mean(d0 ** 2))):.2e}",
            frac_gt_0p01px=f"{float(np.mean(d0 > 0.01)):.2e}")
    assert float(np.mean(d0 > 0.01)) <= far_frac


def test_run_pair_end_to_end_vs_oracle(pivx, orc):
    """1024x768 Lamb-Oseen pair through the sample Parameter.dat pipeline
    (grid 1 64x64 at 32 x2, grid 2 32x32 at 16 x2, 5x5 median x2, fill, 3x3
    smoothing), GPU run_pair against the FP64 oracle's run_pair on the same
    frames.  Bar: final-field RMS difference <= 1e-3 px, at most 0.5% of the
    vectors farther apart than 0.01 px, at most 0.5% of the validity flags
    different."""
    nx, ny = 1024, 768
    sp_ = pivx.synth_params(nx, ny, seed=77, flow=1, a0=150.0, a1=6.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    p = pivx.init_params(nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2, ws_x2=32,
                         ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2)
    fp = pivx.FilterParams(2, 5, 2.0, 3.0, 1)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.set_graph_mode(True)
        x.run_pair(fp, fp)
        gpu = _pair_fields(x)
    ref = orc.run_pair(a, b, p, fp, fp, prec=64, threads=THREADS)
    _compare_fields("end to end 1024x768", gpu, ref, rms_tol=1e-3, far_frac=5e-3, valid_frac=5e-3)


def test_config2_full_size_vs_oracle(pivx, orc):
    """The 6600x4400 config-2 pair against the FP64 oracle: stage-wise (the
    first grid-1 pass, and one deformed grid-2 pass fed the GPU's own
    predictor and deformed frame; bar as for every pass: indices, order and
    validity bit-exact outside ties, sub-pixel within 1e-3 px) and end to end
    (run_pair against the oracle's run_pair; bar as in the mid-size test)."""
    p = pivx.init_params(nx=6600, ny=4400, nsec_cc=1, ws_x1=64, ws_y1=64, spacing_x1=64, spacing_y1=64, npass1=3,
                         ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=3, npeak_save=4,
                         max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                         max_dispy_percent2=0.4)
    fp = pivx.FilterParams(npass_median=2, min_nneigh=5, residual_thr=2.0, insert_thr=2.0, npass_smooth=1)
    sp_ = pivx.synth_params(6600, 4400, seed=2024, density=0.05, flow=1, a0=300.0, a1=8.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    with pivx.Xcorr2dMulti(p) as x:
        g1 = geom_dict(x.geometry(1))
        g2 = geom_dict(x.geometry(2))
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu_g1 = gpu_pass(x, 1)
        x.set_graph_mode(True)
        x.run_pair(fp, fp)
        gpu = _pair_fields(x)
        # one more grid-2 pass from the pipeline's final field
        pred = gpu["filt2"][:2]
        x.set_graph_mode(False)
        x.extend_filtered_vecfield_p2()
        x.Calc_cc_map_p2()
        gpu_g2 = gpu_pass(x, 2)
        d2 = x.deformed(2)
    ref1 = orc.xcorr_pass(a, f2u8=b, geom=g1, ws=(64, 64), sp=(64, 64), K=4, tie_rel=tie_margin(64, 64),
                          threads=THREADS)
    check_pass("config2 grid-1 pass 0", gpu_g1, ref1, (64, 64), 4)
    ref2 = orc.xcorr_pass(a, f2def=d2.astype(np.float64), geom=g2, ws=(32, 32), sp=(16, 16), K=4,
                          pred=(pred[0].astype(np.float64), pred[1].astype(np.float64)), tie_rel=tie_margin(32, 32),
                          threads=THREADS)
    check_pass("config2 grid-2 deformed pass", gpu_g2, ref2, (32, 32), 4)
    ref = orc.run_pair(a, b, p, fp, fp, prec=64, threads=THREADS)
    _compare_fields("end to end 6600x4400", gpu, ref, rms_tol=1e-3, far_frac=5e-3, valid_frac=5e-3)


@pytest.mark.parametrize("ws", [(32, 32), (64, 64), (16, 16), (48, 24), (24, 24)])
def test_delta_windows_zero_neighbours(pivx, orc, ws):
    """Windows holding one bright pixel at their first pixel in both frames
    (no window overlap): the correlation is a delta at zero shift with
    neighbours exactly 0 in the register FFTs (radices 2, 3 and 4: every
    nonzero value meets only trivial twiddles, every other sum is of zeros), so the
    3-point fit takes log(0) = -inf, the estimate is NaN and the NaN removal
    of kernel_remove_nan (cu:457-471) leaves an invalid vector with zero
    displacement.  Every third window is left empty (all zero): no local
    maximum.  GPU and FP64 oracle agree exactly on
    candidate counts, indices, values and validity."""
    nx, ny = 512, 384
    p = pivx.init_params(nx=nx, ny=ny, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=ws[0], spacing_y1=ws[1], npass1=1,
                         npass2=0, max_dispx_percent1=0.4, max_dispy_percent1=0.4)
    with pivx.Xcorr2dMulti(p) as x:
        g = geom_dict(x.geometry(1))
        a = np.zeros((ny, nx), np.uint8)
        b = np.zeros((ny, nx), np.uint8)
        for iy in range(g["ny_vec"]):
            for ix in range(g["nx_vec"]):
                xi, yi = ix * ws[0] - g["left_x"], iy * ws[1] - g["left_y"]
                if 0 <= xi < nx and 0 <= yi < ny and (ix + iy) % 3:
                    a[yi, xi] = 200
                    b[yi, xi] = 100
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=ws, sp=ws, K=4, tie_rel=tie_margin(*ws))
    delta = ref["ncand"] == 1
    _report(f"delta windows {ws[0]}x{ws[1]}", windows=len(delta), delta=int(delta.sum()),
            empty=int((ref["ncand"] == 0).sum()))
    assert delta.sum() > 0 and (ref["ncand"] == 0).sum() > 0
    np.testing.assert_array_equal(gpu["nc"], ref["ncand"])
    np.testing.assert_array_equal(gpu["idx"][delta, 0], ref["pk_idx"][delta, 0])
    np.testing.assert_array_equal(gpu["idx"][delta, 0], (ws[1] // 2) * ws[0] + ws[0] // 2)
    np.testing.assert_array_equal(gpu["val"][delta, 0], np.float32(200 * 100))
    np.testing.assert_array_equal(gpu["val"][delta, 0], ref["pk_val"][delta, 0])
    np.testing.assert_array_equal(gpu["v"], ref["valid"])
    assert not gpu["v"].any()
    np.testing.assert_array_equal(gpu["dx"][delta, 0], 0.0)
    np.testing.assert_array_equal(gpu["dy"][delta, 0], 0.0)
