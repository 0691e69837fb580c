"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bar (written per assertion below):
  * integer peak index, candidate order and validity flags bit-exact, except
    in windows the FP64 oracle flags as ties (two candidate values, or a
    candidate and its neighbour, closer than tie_margin * peak);
  * sub-pixel displacements within 1e-3 px, except in windows the oracle
    flags as ill-conditioned fits (nearly flat peaks, where plane values
    moved by one FP32 rounding per FFT stage, 2^-23 * log2(N) * peak, move
    the 3-point fit by more than 1e-3 px);
  * at most 3% of the windows excluded by either flag;
  * FP32 element-wise stages (deformation, upsampling, median test, fill,
    smoothing) bit-exact against the oracle's FP32 transcription on identical
    inputs; the deformed frame within 0.05 grey levels of the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import config1_pair, config1_params, geom_dict, tie_margin

pytestmark = pytest.mark.gpu

SUBPIX_TOL = 1e-3


def _report(name, **kw):
    print(f"[parity] {name}: " + ", ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def check_pass(name, gpu, ref, ws, K, min_ok=0.97):
    """gpu: dict(val, idx, dx, dy, v); ref: oracle xcorr_pass output."""
    ok = ref["tie"] == 0       # neither flag: everything checked
    ordered = ref["tie"] != 1  # no value tie: indices, order and validity checked
    frac_ok = float(ok.mean())
    bad_idx = int((gpu["idx"][ordered] != ref["pk_idx"][ordered]).any(axis=1).sum())
    bad_valid = int((gpu["v"][ordered] != ref["valid"][ordered]).any(axis=1).sum())
    both = ok[:, None] & (gpu["v"] == 1) & (ref["valid"] == 1)
    ex = float(np.abs(gpu["dx"] - ref["dx"])[both].max()) if both.any() else 0.0
    ey = float(np.abs(gpu["dy"] - ref["dy"])[both].max()) if both.any() else 0.0
    peak = max(float(ref["pk_val"].max()), 1.0)
    relv = float((np.abs(gpu["val"] - ref["pk_val"]) / peak)[ok].max()) if ok.any() else 0.0
    _report(name, windows=len(ok), tie_excluded=int((ref["tie"] == 1).sum()),
            flat_fit_excluded=int((ref["tie"] == 2).sum()), idx_mismatch=bad_idx, valid_mismatch=bad_valid,
            max_subpix_err=f"{max(ex, ey):.2e}", max_rel_peak_err=f"{relv:.2e}")
    assert frac_ok >= min_ok, f"too many tie-excluded windows ({1 - frac_ok:.1%})"
    assert bad_idx == 0, "peak indices differ outside ties"
    assert bad_valid == 0, "validity flags differ outside ties"
    assert max(ex, ey) <= SUBPIX_TOL
    assert relv <= tie_margin(*ws)


def gpu_pass(x, grid):
    val, idx, nc = x.peaks(grid)
    dx, dy, v = x.raw(grid)
    return dict(val=val, idx=idx, dx=dx, dy=dy, v=v, nc=nc)


def test_config1_pass0_matches_oracle(pivx, orc):
    """512x512, 32x32 @16, top-4, undeformed pass."""
    a, b = config1_pair(pivx)
    p = config1_params(pivx)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
        g = geom_dict(x.geometry(1))
    assert g["nx_vec"] == 32 and g["ny_vec"] == 32
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=(32, 32), sp=(16, 16), K=4, tie_rel=tie_margin(32, 32))
    check_pass("config1 pass0", gpu, ref, (32, 32), 4)
    # uncapped candidate counts agree wherever the window is not tie-flagged
    ok = ref["tie"] == 0
    assert (gpu["nc"][ok] == ref["ncand"][ok]).mean() > 0.98
    # interior vectors recover the imposed (3.3, -1.7) displacement
    gx, gy = np.meshgrid(np.arange(32) * 16 + 8, np.arange(32) * 16 + 8)
    interior = ((gx > 24) & (gx < 488) & (gy > 24) & (gy < 488)).ravel()
    good = interior & (gpu["v"][:, 0] == 1)
    assert np.median(np.abs(gpu["dx"][good, 0] - 3.3)) < 0.1
    assert np.median(np.abs(gpu["dy"][good, 0] + 1.7)) < 0.1


def test_deformation_bitexact_and_deformed_pass(pivx, orc):
    """Kernel_Window_Deformation_2d against the FP32 transcription (bit-exact)
    and the FP64 oracle; then the deformed correlation pass on identical
    inputs."""
    a, b = config1_pair(pivx, seed=3)
    p = config1_params(pivx, npass1=2)
    rng = np.random.default_rng(7)
    with pivx.Xcorr2dMulti(p) as x:
        g = geom_dict(x.geometry(1))
        nt = g["nx_vec"] * g["ny_vec"]
        fx = (3.3 + 0.3 * rng.standard_normal(nt)).astype(np.float32)
        fy = (-1.7 + 0.3 * rng.standard_normal(nt)).astype(np.float32)
        x.prep_img_data(a, b)
        x.set_filtered(1, fx, fy)
        x.extend_filtered_vecfield_p1()
        x.Calc_cc_map_p1(1)
        d = x.deformed(1)
        gpu = gpu_pass(x, 1)
    ex, ey = orc.extend(fx.reshape(32, 32), fy.reshape(32, 32))
    d32 = orc.deform(b, g, (32, 32), (16, 16), ex, ey, prec=32)
    d64 = orc.deform(b, g, (32, 32), (16, 16), ex, ey, prec=64)
    _report("deformation", shape=d.shape, max_diff_f32=float(np.abs(d - d32).max()),
            max_diff_f64=f"{float(np.abs(d - d64).max()):.2e}")
    np.testing.assert_array_equal(d, d32.astype(np.float32))
    # FP32 sample positions near x ~ 500 carry ~3e-5 px of rounding; times
    # intensity gradients up to ~255 grey levels/px that is ~0.01-0.02.
    assert np.abs(d - d64).max() < 0.05
    ref = orc.xcorr_pass(a, f2def=d.astype(np.float64), geom=g, ws=(32, 32), sp=(16, 16), K=4,
                         pred=(fx.astype(np.float64), fy.astype(np.float64)), tie_rel=tie_margin(32, 32))
    check_pass("config1 deformed pass", gpu, ref, (32, 32), 4)


@pytest.mark.parametrize("shift", [(20, -18), (-17, 23), (5, -3)])
def test_deformation_out_of_buffer_taps(pivx, shift):
    """The count of deformation taps beyond the reference's padded buffer
    (aug coordinates outside [-hfws, n_aug + hfws - 1)).  An integer uniform
    field shifts every aug pixel exactly (the bilinear weights of a spacing-16
    grid are exact binary fractions summing to 1), so the count has a closed
    form; shifts beyond the half window exercise the exact per-pixel test,
    the small shift the per-cell bound that skips it."""
    a, b = config1_pair(pivx, seed=4)
    p = config1_params(pivx, npass1=2)
    with pivx.Xcorr2dMulti(p) as x:
        g = geom_dict(x.geometry(1))
        nt = g["nx_vec"] * g["ny_vec"]
        x.prep_img_data(a, b)
        x.set_filtered(1, np.full(nt, shift[0], np.float32), np.full(nt, shift[1], np.float32))
        x.extend_filtered_vecfield_p1()
        x.Calc_cc_map_p1(1)
        got = x.stats()["oob_taps"][0]
    hfws = 16
    jx = np.arange(g["nx_aug"]) + shift[0]
    jy = np.arange(g["ny_aug"]) + shift[1]
    out_x = (jx + hfws < 0) | (jx + hfws >= g["nx_aug"] + 2 * hfws - 1)
    out_y = (jy + hfws < 0) | (jy + hfws >= g["ny_aug"] + 2 * hfws - 1)
    want = int((out_x[None, :] | out_y[:, None]).sum())
    _report("out-of-buffer taps", shift=shift, got=got, want=want)
    assert got == want
    if max(abs(shift[0]), abs(shift[1])) < hfws - 2:
        assert want == 0


@pytest.mark.parametrize("ws,big", [(32, (13.0, 24.0)), (64, (29.0, 45.0))])
def test_deformation_mixed_cells_bitexact(pivx, orc, ws, big):
    """A field of small vectors (every tap inside frame 2's zero border, read
    without range tests) with scattered vectors beyond hfws - 2 (cells whose
    taps may leave the reference's padded buffer, read with the checked
    path; the corner vector sends taps beyond it), so that both kinds of cell
    share waves.  The deformed plane equals the FP32 transcription bit for
    bit; at 64x64 windows on a spacing of 64 the deformation fused into the
    correlation kernel gives the same correlation results and out-of-buffer
    count as the deformed plane."""
    a, b = config1_pair(pivx, seed=5)
    p = config1_params(pivx, ws_x1=ws, ws_y1=ws, spacing_x1=ws // 2 if ws == 32 else ws,
                       spacing_y1=ws // 2 if ws == 32 else ws, npass1=2)
    rng = np.random.default_rng(29)
    res = {}
    for mode in ((0,) if ws == 32 else (0, 2)):
        with pivx.Xcorr2dMulti(p) as x:
            x.set_fused_deform(mode)
            g = geom_dict(x.geometry(1))
            nvx, nvy = g["nx_vec"], g["ny_vec"]
            nt = nvx * nvy
            if mode == 0:
                fx = (2.0 + rng.standard_normal(nt)).astype(np.float32)
                fy = (-1.0 + rng.standard_normal(nt)).astype(np.float32)
                pick = rng.random(nt) < 0.15
                mag = rng.uniform(big[0], big[1], size=(2, nt)) * rng.choice([-1.0, 1.0], size=(2, nt))
                fx[pick] = mag[0, pick]
                fy[pick] = mag[1, pick]
                fx[0] = fy[0] = -big[1]
            x.prep_img_data(a, b)
            x.set_filtered(1, fx, fy)
            x.extend_filtered_vecfield_p1()
            x.Calc_cc_map_p1(1)
            res[mode] = (x.stats()["oob_taps"][0], x.raw(1) + x.peaks(1))
            if mode == 0:
                d = x.deformed(1)
    ex, ey = orc.extend(fx.reshape(nvy, nvx), fy.reshape(nvy, nvx))
    sp = ws // 2 if ws == 32 else ws
    d32 = orc.deform(b, g, (ws, ws), (sp, sp), ex, ey, prec=32)
    _report("deformation mixed cells", ws=ws, large=int(pick.sum()), oob=res[0][0],
            max_diff_f32=float(np.abs(d - d32).max()))
    np.testing.assert_array_equal(d, d32.astype(np.float32))
    assert res[0][0] > 0
    if 2 in res:
        assert res[2][0] == res[0][0]
        for u, w in zip(res[0][1], res[2][1]):
            np.testing.assert_array_equal(u, w)


@pytest.mark.parametrize("shift", [(40, -36), (-35, 45), (10, -6)])
def test_fused_deformation_matches_plane(pivx, shift):
    """Deformation fused into the correlation kernel (grid-1 64x64 windows at
    spacing 64, pivx_set_fused_deform(2)) against the deformed-plane path
    (mode 0): identical correlation results, and the out-of-buffer tap count
    equals the closed form of an integer shift in both (the weights of a
    spacing-64 grid are exact binary fractions summing to 1)."""
    a, b = config1_pair(pivx, seed=4)
    p = config1_params(pivx, ws_x1=64, ws_y1=64, spacing_x1=64, spacing_y1=64, npass1=2)
    res = {}
    for mode in (0, 2):
        with pivx.Xcorr2dMulti(p) as x:
            x.set_fused_deform(mode)
            g = geom_dict(x.geometry(1))
            nt = g["nx_vec"] * g["ny_vec"]
            x.prep_img_data(a, b)
            x.set_filtered(1, np.full(nt, shift[0], np.float32), np.full(nt, shift[1], np.float32))
            x.extend_filtered_vecfield_p1()
            x.Calc_cc_map_p1(1)
            res[mode] = (x.stats()["oob_taps"][0], x.raw(1) + x.peaks(1))
    hfws = 32
    jx = np.arange(g["nx_aug"]) + shift[0]
    jy = np.arange(g["ny_aug"]) + shift[1]
    out_x = (jx + hfws < 0) | (jx + hfws >= g["nx_aug"] + 2 * hfws - 1)
    out_y = (jy + hfws < 0) | (jy + hfws >= g["ny_aug"] + 2 * hfws - 1)
    want = int((out_x[None, :] | out_y[:, None]).sum())
    _report("fused deformation", shift=shift, oob_plane=res[0][0], oob_fused=res[2][0], want=want)
    assert res[0][0] == want and res[2][0] == want
    for u, w in zip(res[0][1], res[2][1]):
        np.testing.assert_array_equal(u, w)


@pytest.mark.parametrize("ws", [(64, 64), (48, 48), (64, 32), (32, 64), (48, 24)])
def test_fused_deformation_window_shapes(pivx, ws):
    """Every window shape of the wave-per-window kernel with the deformation
    fused in (spacing = window, a random predictor field, 448x352 frames, of
    which the 48-wide windows leave a strip uncovered): raw candidates, peak
    data and out-of-buffer counts identical to the deformed-plane path."""
    a, b = config1_pair(pivx, seed=21, nx=448, ny=352, u=2.1, v=-1.6)
    p = config1_params(pivx, nx=448, ny=352, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=ws[0], spacing_y1=ws[1], npass1=2,
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4)
    rng = np.random.default_rng(23)
    res = {}
    for mode in (0, 2):
        with pivx.Xcorr2dMulti(p) as x:
            x.set_fused_deform(mode)
            g = geom_dict(x.geometry(1))
            nt = g["nx_vec"] * g["ny_vec"]
            if mode == 0:
                fx = (2.1 + 0.7 * rng.standard_normal(nt)).astype(np.float32)
                fy = (-1.6 + 0.7 * rng.standard_normal(nt)).astype(np.float32)
            x.prep_img_data(a, b)
            x.set_filtered(1, fx, fy)
            x.extend_filtered_vecfield_p1()
            x.Calc_cc_map_p1(1)
            res[mode] = (x.stats()["oob_taps"][0], x.raw(1) + x.peaks(1))
    _report("fused deformation shape", ws=ws, windows=nt, valid=int((res[2][1][2][:, 0] == 1).sum()))
    assert res[0][0] == res[2][0]
    for u, w in zip(res[0][1], res[2][1]):
        np.testing.assert_array_equal(u, w)


@pytest.mark.parametrize("ws", [(16, 16), (24, 24), (48, 48), (64, 64), (64, 32), (32, 64), (32, 16), (16, 32),
                                (48, 24)])
def test_fast_window_shapes(pivx, orc, ws):
    """Every compiled register-FFT window shape, undeformed pass, 50% overlap."""
    a, b = config1_pair(pivx, seed=11, nx=384, ny=320, u=2.6, v=1.3)
    sp = (ws[0] // 2, ws[1] // 2)
    p = config1_params(pivx, nx=384, ny=320, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=sp[0], spacing_y1=sp[1],
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
        g = geom_dict(x.geometry(1))
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=ws, sp=sp, K=4, tie_rel=tie_margin(*ws))
    check_pass(f"fast {ws[0]}x{ws[1]}", gpu, ref, ws, 4)


@pytest.mark.parametrize("ws,pct", [((32, 32), (0.15, 0.3)), ((64, 64), (0.3, 0.1)), ((48, 48), (0.05, 0.45)),
                                    ((32, 32), (0.0, 0.4)), ((64, 32), (0.45, 0.0))])
def test_search_border_shapes(pivx, orc, ws, pct):
    """Peak search borders other than the default (cu:596): wide and
    asymmetric ones, rows crossing row 32 of a 64-row window, and an empty
    search region (border ws/2: no candidates anywhere).  Candidate counts
    agree in 98 % of the windows outside ties, as in the default-border test
    (a minor maximum whose neighbour is within rounding of it may count on
    one side only), and are 0 everywhere for an empty region."""
    a, b = config1_pair(pivx, seed=17, nx=384, ny=320, u=1.7, v=-2.4)
    sp = (ws[0] // 2, ws[1] // 2)
    p = config1_params(pivx, nx=384, ny=320, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=sp[0], spacing_y1=sp[1],
                       max_dispx_percent1=pct[0], max_dispy_percent1=pct[1])
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
        g = geom_dict(x.geometry(1))
    assert (g["border_x"], g["border_y"]) == (max(1, ws[0] // 2 - round(ws[0] * pct[0])),
                                              max(1, ws[1] // 2 - round(ws[1] * pct[1])))
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=ws, sp=sp, K=4, tie_rel=tie_margin(*ws))
    check_pass(f"border {g['border_x']},{g['border_y']} at {ws[0]}x{ws[1]}", gpu, ref, ws, 4)
    ok = ref["tie"] == 0
    assert (gpu["nc"][ok] == ref["ncand"][ok]).mean() > 0.98
    if g["border_x"] * 2 >= ws[0] or g["border_y"] * 2 >= ws[1]:
        assert int(np.asarray(gpu["nc"]).sum()) == 0 and (gpu["idx"] == -1).all()


@pytest.mark.parametrize("ws,sp", [((20, 12), (10, 6)), ((36, 28), (18, 14)), ((44, 44), (22, 22)),
                                   ((128, 128), (64, 64)), ((8, 8), (4, 4))])
def test_generic_window_shapes(pivx, orc, ws, sp):
    """Shapes outside the compiled set take the generic kernel (runtime radix
    plan; 128x128 uses global-memory planes, on a larger frame so that the
    excluded fraction is measured over enough windows)."""
    nx, ny = (1024, 768) if ws[0] >= 128 else (320, 256)
    a, b = config1_pair(pivx, seed=5, nx=nx, ny=ny, u=1.4, v=-2.2)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=sp[0], spacing_y1=sp[1],
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
        g = geom_dict(x.geometry(1))
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=ws, sp=sp, K=4, tie_rel=tie_margin(*ws))
    # a 128x128 window holds 16 times the local maxima of a 32x32 one, so
    # near-equal candidate / neighbour values are more frequent (7 of 192
    # windows measured): 95% there
    check_pass(f"generic {ws[0]}x{ws[1]}", gpu, ref, ws, 4, min_ok=0.95 if ws[0] >= 128 else 0.97)


@pytest.mark.parametrize("ws,sp", [((32, 32), (16, 16)), ((64, 64), (32, 32)), ((48, 24), (24, 12))])
def test_noise_windows_near_ties(pivx, orc, ws, sp):
    """Config-1 images with 5% of the windows carrying added N(0, 20) noise in
    both frames (near-equal secondary peaks): outside the oracle's tie flags
    every peak index and validity flag still matches and the sub-pixel
    displacements agree within 1e-3 px."""
    a, b = config1_pair(pivx, seed=31)
    p = config1_params(pivx, ws_x1=ws[0], ws_y1=ws[1], spacing_x1=sp[0], spacing_y1=sp[1])
    rng = np.random.default_rng(32)
    with pivx.Xcorr2dMulti(p) as x:
        g = geom_dict(x.geometry(1))
        nxv, nyv = g["nx_vec"], g["ny_vec"]
        noisy = rng.random(nxv * nyv) < 0.05
        a = a.astype(np.float64)
        b = b.astype(np.float64)
        for w in np.flatnonzero(noisy):
            # window w covers image columns [c - ws/2, c + ws/2) around c = sp/2 + k*sp
            cx = sp[0] // 2 + (w % nxv) * sp[0]
            cy = sp[1] // 2 + (w // nxv) * sp[1]
            ys = slice(max(cy - ws[1] // 2, 0), max(cy + ws[1] // 2, 0))
            xs = slice(max(cx - ws[0] // 2, 0), max(cx + ws[0] // 2, 0))
            a[ys, xs] += rng.normal(0.0, 20.0, a[ys, xs].shape)
            b[ys, xs] += rng.normal(0.0, 20.0, b[ys, xs].shape)
        a = np.clip(np.rint(a), 0, 255).astype(np.uint8)
        b = np.clip(np.rint(b), 0, 255).astype(np.uint8)
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=ws, sp=sp, K=4, tie_rel=tie_margin(*ws))
    check_pass(f"noise windows {ws[0]}x{ws[1]} ({int(noisy.sum())} noisy)", gpu, ref, ws, 4)


def test_uneven_batch_nonsquare_windows(pivx, orc):
    """Pairs of uneven size (1024x1024, 2048x1536, 1280x720) with random
    Fourier-mode flows (max 10 px, densities 0.02-0.08) through the two-grid
    pipeline with non-square windows: 64x32 at 32x16, then 32x16 at 16x8.
    Each pair runs on its own engine context, all contexts enqueued back to
    back; every result equals the same pair run alone, the grid-2 field follows
    the analytic flow, and the first grid-1 pass of the smallest pair matches
    the oracle."""
    sizes = [(1024, 1024, 0.05), (2048, 1536, 0.02), (1280, 720, 0.08)]
    fp = pivx.FilterParams(npass_median=2, min_nneigh=5, residual_thr=2.0, insert_thr=2.0, npass_smooth=1)

    def params(nx, ny):
        return config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=32, spacing_x1=32, spacing_y1=16, npass1=2,
                              ws_x2=32, ws_y2=16, spacing_x2=16, spacing_y2=8, npass2=2,
                              max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                              max_dispy_percent2=0.4)

    pairs = []
    for i, (nx, ny, dens) in enumerate(sizes):
        sp_ = pivx.synth_params(nx, ny, seed=700 + i, density=dens, flow=2, a0=10.0)
        pairs.append((sp_, *pivx.synth_pair(sp_)))

    def fields(x):
        return [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]

    alone = []
    for sp_, a, b in pairs:
        with pivx.Xcorr2dMulti(params(sp_.nx, sp_.ny)) as x:
            x.prep_img_data(a, b)
            x.set_graph_mode(True)
            x.run_pair(fp, fp)
            alone.append((fields(x), x.grid_coords(2)))
    ctxs = [pivx.Xcorr2dMulti(params(sp_.nx, sp_.ny)) for sp_, _, _ in pairs]
    try:
        for x, (sp_, a, b) in zip(ctxs, pairs):
            x.prep_img_data(a, b)
            x.set_graph_mode(True)
        for x in ctxs:
            x.run_pair(fp, fp)
        batch = [fields(x) for x in ctxs]
    finally:
        for x in ctxs:
            x.close()
    for (ref, (gx, gy)), got, (sp_, a, b) in zip(alone, batch, pairs):
        for A, B in zip(ref, got):
            for u, w in zip(A, B):
                np.testing.assert_array_equal(u, w)
        u, v = pivx.synth_flow(sp_, gx, gy)
        fx, fy, fv = ref[3]
        inner = (gx > 48) & (gx < sp_.nx - 48) & (gy > 48) & (gy < sp_.ny - 48)
        rms = float(np.sqrt(np.mean((fx[inner] - u[inner]) ** 2 + (fy[inner] - v[inner]) ** 2)))
        _report("uneven batch", size=(sp_.nx, sp_.ny), density=sp_.density, rms_px=f"{rms:.4f}",
                valid=f"{float(np.mean(fv[inner])):.4f}")
        assert rms < 0.25
        assert float(np.mean(fv[inner])) > 0.99
    sp_, a, b = pairs[2]
    with pivx.Xcorr2dMulti(params(sp_.nx, sp_.ny)) as x:
        x.prep_img_data(a, b)
        x.Calc_cc_map_p1(0)
        gpu = gpu_pass(x, 1)
        g = geom_dict(x.geometry(1))
    ref = orc.xcorr_pass(a, f2u8=b, geom=g, ws=(64, 32), sp=(32, 16), K=4, tie_rel=tie_margin(64, 32))
    check_pass("uneven batch 1280x720 grid-1 pass 0 (64x32)", gpu, ref, (64, 32), 4)


def _random_candidates(rng, nt, K, outlier_frac=0.1):
    dx = (2.0 + 0.2 * rng.standard_normal((nt, K))).astype(np.float32)
    dy = (-1.0 + 0.2 * rng.standard_normal((nt, K))).astype(np.float32)
    out = rng.random(nt) < outlier_frac
    dx[out, 0] += rng.uniform(-8, 8, out.sum()).astype(np.float32)
    dy[out, 0] += rng.uniform(-8, 8, out.sum()).astype(np.float32)
    v = (rng.random((nt, K)) > 0.15).astype(np.uint8)
    dx[v == 0] = 0
    dy[v == 0] = 0
    return dx, dy, v


@pytest.mark.parametrize("win,npm,nps,fill", [((5, 5), 2, 1, 1), ((3, 3), 1, 2, 1), ((7, 5), 3, 0, 0),
                                              ((9, 9), 1, 1, 1), ((4, 4), 2, 1, 1), ((7, 7), 3, 3, 1),
                                              ((5, 5), 0, 1, 1), ((5, 5), 4, 2, 0)])
def test_validation_bitexact(pivx, orc, win, npm, nps, fill):
    """MedianFilter_Smoothing_p2 on injected candidates: median passes, fill,
    smoothing, bit-exact against the FP32 transcription."""
    rng = np.random.default_rng(win[0] * 10 + npm)
    p = config1_params(pivx, nx=640, ny=480, ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=1,
                       wx2_median=win[0], wy2_median=win[1], wx2_smooth=3, wy2_smooth=5)
    with pivx.Xcorr2dMulti(p) as x:
        nxv, nyv = x.dims(2)
        nt = nxv * nyv
        dx, dy, v = _random_candidates(rng, nt, 4)
        v[rng.random(nt) < 0.02, :] = 0  # some vectors without any valid candidate
        x.set_raw(2, dx, dy, v)
        x.MedianFilter_Smoothing_p2(npm, 5, 2.0, 2.0, nps, fill)
        gx, gy, gv = x.filtered(2)
        st = x.stats()
    K = 4
    cx, cy, cv = dx.astype(np.float64).ravel(), dy.astype(np.float64).ravel(), v.ravel()
    sx, sy, sv = dx[:, 0].astype(np.float64), dy[:, 0].astype(np.float64), v[:, 0].copy()
    for _ in range(npm):
        sx, sy, sv, _m = orc.median_pass((sx, sy, sv), (cx, cy, cv), K, (nxv, nyv), win, 5, 2.0, prec=32)
    sweeps = None
    if fill:
        sx, sy, sv, sweeps = orc.fill(sx, sy, sv, (nxv, nyv), prec=32)
    for _ in range(nps):
        sx, sy = orc.smooth(sx, sy, (nxv, nyv), (3, 5), prec=32)
    _report(f"validation {win} npm={npm} nps={nps} fill={fill}", vectors=nt, invalid=int((sv == 0).sum()),
            max_diff=float(max(np.abs(gx - sx).max(), np.abs(gy - sy).max())), sweeps_gpu=st["fill_sweeps"][1],
            sweeps_oracle=sweeps)
    np.testing.assert_array_equal(gv, sv)
    np.testing.assert_array_equal(gx, sx.astype(np.float32))
    np.testing.assert_array_equal(gy, sy.astype(np.float32))
    if fill:
        assert st["fill_sweeps"][1] == sweeps


def test_upsampling_bitexact(pivx, orc):
    p = config1_params(pivx, nx=1000, ny=700, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass2=1,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16)
    rng = np.random.default_rng(3)
    with pivx.Xcorr2dMulti(p) as x:
        n1 = x.dims(1)
        n2 = x.dims(2)
        fx = rng.standard_normal(n1[0] * n1[1]).astype(np.float32)
        fy = rng.standard_normal(n1[0] * n1[1]).astype(np.float32)
        x.set_filtered(1, fx, fy)
        x.extend_filtered_vecfield_p1()
        x.Upsampling_vector_field()
        ux, uy, _ = x.filtered(2)
    ex, ey = orc.extend(fx.reshape(n1[1], n1[0]), fy.reshape(n1[1], n1[0]))
    rx, ry = orc.upsample(ex, ey, (32, 32), (16, 16), n2, prec=32)
    np.testing.assert_array_equal(ux, rx.astype(np.float32))
    np.testing.assert_array_equal(uy, ry.astype(np.float32))


def test_two_grid_stagewise(pivx, orc):
    """Two-grid, multi-pass pipeline on a Lamb-Oseen flow; at every pass the
    oracle is fed the GPU's own inputs of that stage (predictor, deformed
    frame, raw candidates) so stage errors do not compound."""
    nx, ny = 768, 512
    sp_ = pivx.synth_params(nx, ny, seed=21, flow=1, a0=120.0, a1=6.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2,
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                       max_dispy_percent2=0.4)
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        geo = {1: geom_dict(x.geometry(1)), 2: geom_dict(x.geometry(2))}
        wss = {1: (64, 64), 2: (32, 32)}
        sps = {1: (32, 32), 2: (16, 16)}
        for grid, npass in ((1, 2), (2, 2)):
            g = geo[grid]
            nvec = (g["nx_vec"], g["ny_vec"])
            if grid == 2:
                x.extend_filtered_vecfield_p1()
                x.Upsampling_vector_field()
            for j in range(npass):
                if grid == 1 and j > 0:
                    x.extend_filtered_vecfield_p1()
                if grid == 2:
                    x.extend_filtered_vecfield_p2()
                pred = x.filtered(grid)[:2] if (grid == 2 or j > 0) else (np.zeros(nvec[0] * nvec[1]),) * 2
                if grid == 1:
                    x.Calc_cc_map_p1(j)
                else:
                    x.Calc_cc_map_p2()
                gpu = gpu_pass(x, grid)
                deformed = grid == 2 or j > 0
                kw = dict(geom=g, ws=wss[grid], sp=sps[grid], K=4, tie_rel=tie_margin(*wss[grid]),
                          pred=(pred[0].astype(np.float64), pred[1].astype(np.float64)))
                if deformed:
                    d = x.deformed(grid)
                    ex, ey = orc.extend(pred[0].reshape(nvec[1], nvec[0]), pred[1].reshape(nvec[1], nvec[0]))
                    d32 = orc.deform(b, g, wss[grid], sps[grid], ex, ey, prec=32)
                    np.testing.assert_array_equal(d, d32.astype(np.float32))
                    ref = orc.xcorr_pass(a, f2def=d.astype(np.float64), **kw)
                else:
                    ref = orc.xcorr_pass(a, f2u8=b, **kw)
                check_pass(f"two-grid g{grid} pass{j}", gpu, ref, wss[grid], 4)
                # validation on the GPU's raw candidates
                if grid == 1:
                    x.MedianFilter_Smoothing_p1(2, 5, 2.0, 2.0, 1)
                else:
                    x.MedianFilter_Smoothing_p2(2, 5, 2.0, 2.0, 1, 1)
                fx, fy, fv = x.filtered(grid)
                cx, cy, cv = (gpu["dx"].astype(np.float64).ravel(), gpu["dy"].astype(np.float64).ravel(),
                              gpu["v"].ravel())
                sx, sy, sv = gpu["dx"][:, 0].astype(np.float64), gpu["dy"][:, 0].astype(np.float64), gpu["v"][:, 0]
                for _ in range(2):
                    sx, sy, sv, _m = orc.median_pass((sx, sy, sv), (cx, cy, cv), 4, nvec, (5, 5), 5, 2.0, prec=32)
                sx, sy, sv, _s = orc.fill(sx, sy, sv, nvec, prec=32)
                sx, sy = orc.smooth(sx, sy, nvec, (3, 3), prec=32)
                np.testing.assert_array_equal(fv, sv)
                np.testing.assert_array_equal(fx, sx.astype(np.float32))
                np.testing.assert_array_equal(fy, sy.astype(np.float32))
        del fp


def test_config2_full_size_properties(pivx):
    """The 6600x4400 Lamb-Oseen pair (vortex core 300 px, peak swirl 8 px,
    shear 0.002) through the whole two-grid pipeline: grid 1 64x64 at 64 x3,
    grid 2 32x32 at 16 x3, top-4, 5x5 median test x2, fill, 3x3 smoothing.
    Size-independent properties: the grid-2 field against the analytic flow,
    run-to-run bit identity, and bit identity of pairs run concurrently on
    three engine contexts (own streams and pair graphs, enqueued back to back)
    with the one-context result."""
    p = pivx.init_params(nx=6600, ny=4400, nsec_cc=1, ws_x1=64, ws_y1=64, spacing_x1=64, spacing_y1=64, npass1=3,
                         ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=3, npeak_save=4,
                         max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                         max_dispy_percent2=0.4)
    fp = pivx.FilterParams(npass_median=2, min_nneigh=5, residual_thr=2.0, insert_thr=2.0, npass_smooth=1)
    sp_ = pivx.synth_params(6600, 4400, seed=2024, density=0.05, flow=1, a0=300.0, a1=8.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)

    def fields(x):
        return [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]

    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.set_graph_mode(True)
        x.run_pair(fp, fp)
        ref = fields(x)
        x.run_pair(fp, fp)
        again = fields(x)
        gx, gy = x.grid_coords(2)
        st = x.stats()
        pipes = [pivx.Xcorr2dMulti(p) for _ in range(3)]
        try:
            for q in pipes:
                q.prep_img_data(a, b)
                q.set_graph_mode(True)
            for _ in range(2):
                for q in pipes:
                    q.run_pair(fp, fp)
            conc = [fields(q) for q in pipes]
        finally:
            for q in pipes:
                q.close()
    for other in [again] + conc:
        for A, B in zip(ref, other):
            for u, w in zip(A, B):
                np.testing.assert_array_equal(u, w)
    u, v = pivx.synth_flow(sp_, gx, gy)
    fx, fy, fv = ref[3]
    inner = (gx > 64) & (gx < 6600 - 64) & (gy > 64) & (gy < 4400 - 64)
    rms = float(np.sqrt(np.mean((fx[inner] - u[inner]) ** 2 + (fy[inner] - v[inner]) ** 2)))
    valid = float(np.mean(fv[inner]))
    _report("config2 full size", rms_px=f"{rms:.4f}", valid=f"{valid:.5f}", outliers=st["outliers"],
            overflow_windows=st["overflow_windows"])
    assert rms < 0.06
    assert valid > 0.999


@pytest.mark.parametrize("sp1", [32, 64])
def test_run_pair_matches_stepwise_and_is_deterministic(pivx, sp1):
    """pivx_run_pair reproduces the reference's main-loop sequence exactly
    (same result as calling the per-pass entry points) and is bit-identical
    across runs.  With grid-1 windows at spacing 64 (no overlap) run_pair
    deforms frame 2 inside the correlation kernel while the per-pass entry
    points write the deformed plane: the fields and the out-of-buffer tap
    counts must still agree bit for bit."""
    nx, ny = 640, 480
    sp_ = pivx.synth_params(nx, ny, seed=4, flow=1, a0=100.0, a1=5.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=sp1, spacing_y1=sp1, npass1=2,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2,
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                       max_dispy_percent2=0.4)
    f1 = pivx.FilterParams(2, 5, 2.0, 3.0, 1)
    f2 = pivx.FilterParams(2, 5, 2.0, 3.0, 1)
    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.run_pair(f1, f2)
        r1 = [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]
        oob1 = list(x.stats()["oob_taps"])
        x.run_pair(f1, f2)
        r2 = [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]
    with pivx.Xcorr2dMulti(p) as y:
        y.prep_img_data(a, b)
        for j in range(2):
            if j > 0:
                y.extend_filtered_vecfield_p1()
            y.Calc_cc_map_p1(j)
            y.MedianFilter_Smoothing_p1(2, 5, 2.0, 3.0, 1)
        y.extend_filtered_vecfield_p1()
        y.Upsampling_vector_field()
        for j in range(2):
            y.extend_filtered_vecfield_p2()
            y.Calc_cc_map_p2()
            y.MedianFilter_Smoothing_p2(2, 5, 2.0, 3.0, 1, 1)
        s = [y.raw(1), y.filtered(1), y.raw(2), y.filtered(2)]
        y.MedianFilter_Smoothing_p2(2, 5, 2.0, 3.0, 0, 0)
        s.append(y.filtered(2))
        oob_s = list(y.stats()["oob_taps"])
    _report("run_pair vs per-pass", grid1_spacing=sp1, oob_taps=oob1)
    for A, B, S in zip(r1, r2, s):
        for u, w, z in zip(A, B, S):
            np.testing.assert_array_equal(u, w)
            np.testing.assert_array_equal(u, z)
    assert oob1 == oob_s


def test_graph_replays_match_eager(pivx):
    """Graph mode: the captured per-pair sequence replayed several times, with
    eager work and host synchronisation between replays, gives the eager
    results bit for bit every time."""
    nx, ny = 640, 480
    sp_ = pivx.synth_params(nx, ny, seed=8, flow=1, a0=100.0, a1=5.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    a2, b2 = pivx.synth_pair(pivx.synth_params(nx, ny, seed=9, flow=0, a0=1.0, a1=2.0))
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2,
                       max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4,
                       max_dispy_percent2=0.4)
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)

    def grab(x):
        return [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]

    with pivx.Xcorr2dMulti(p) as x:
        x.prep_img_data(a, b)
        x.run_pair(fp, fp)
        ref = grab(x)
        x.prep_img_data(a2, b2)
        x.run_pair(fp, fp)
        ref2 = grab(x)
        x.set_graph_mode(True)
        for rep in range(3):
            x.prep_img_data(a, b)
            x.run_pair(fp, fp)
            got = grab(x)
            for A, B in zip(ref, got):
                for u, w in zip(A, B):
                    np.testing.assert_array_equal(u, w)
            x.set_graph_mode(False)
            x.prep_img_data(a2, b2)
            x.run_pair(fp, fp)            # eager work between replays
            x.set_graph_mode(True)
            x.prep_img_data(a2, b2)
            x.run_pair(fp, fp)
            got2 = grab(x)
            for A, B in zip(ref2, got2):
                for u, w in zip(A, B):
                    np.testing.assert_array_equal(u, w)
    _report("graph replay", replays=6)


# ---------------------------------------------------------------- MHE enhancement
@pytest.mark.parametrize("nx,ny,ws,thr", [(6600, 4400, 40, 0.91), (1000, 700, 48, 0.5), (257, 129, 16, 0.0),
                                          (300, 200, 200, 0.3), (640, 448, 40, 0.91), (333, 222, 10, 0.6),
                                          (128, 96, 7, 0.2), (97, 61, 4, 0.91)])
def test_mhe_bitexact(pivx, orc, nx, ny, ws, thr):
    """mhe_gpu::processsing on the GPU against the oracle's restatement:
    integer work, bit-exact; covers divisible and non-divisible sizes, a
    single-tile row, and windows not divisible by 4 (two-kernel path)."""
    sp_ = pivx.synth_params(nx, ny, seed=nx + ws, flow=0, a0=1.0, a1=0.5)
    a, _ = pivx.synth_pair(sp_)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=32, ws_y1=32, spacing_x1=32, spacing_y1=32, npass2=0)
    with pivx.Xcorr2dMulti(p) as x:
        x.set_mhe(True, ws, thr)
        got = x.mhe_enhance(a)
        got2 = x.mhe_enhance(255 - a)  # second frame through the same buffers
    np.testing.assert_array_equal(got, orc.mhe(a, ws, thr))
    np.testing.assert_array_equal(got2, orc.mhe(255 - a, ws, thr))
    _report(f"mhe {nx}x{ny} ws={ws} thr={thr}", pixels=nx * ny, changed=int((got != a).sum()))


@pytest.mark.parametrize("nx,ny,ws,const", [(1000, 1000, 1000, 77), (2048, 1100, 1024, None)])
def test_mhe_large_windows(pivx, orc, nx, ny, ws, const):
    """Windows of 1000 and 1024 px (several load batches per lane in the
    histogram kernel): a constant frame, whose every pixel falls on one grey
    level, and a particle frame with two tiles; bit-exact against the
    oracle."""
    if const is None:
        a, _ = pivx.synth_pair(pivx.synth_params(nx, ny, seed=5, flow=0, a0=1.0, a1=0.5))
    else:
        a = np.full((ny, nx), const, np.uint8)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=32, ws_y1=32, spacing_x1=32, spacing_y1=32, npass2=0)
    with pivx.Xcorr2dMulti(p) as x:
        x.set_mhe(True, ws, 0.5)
        got = x.mhe_enhance(a)
    np.testing.assert_array_equal(got, orc.mhe(a, ws, 0.5))
    _report(f"mhe limits {nx}x{ny} ws={ws}", changed=int((got != a).sum()))


@pytest.mark.parametrize("sp1", [32, 64])
def test_mhe_run_pair_equals_enhanced_frames(pivx, orc, sp1):
    """With MHE switched on, run_pair equals run_pair on frames the oracle
    enhanced beforehand (main.cpp:267-283 order: enhance, then prep).  The
    enhanced frame 2 is written into its zero-bordered buffer; at a grid-1
    spacing of 64 run_pair also deforms it inside the correlation kernel."""
    nx, ny = 768, 512
    sp_ = pivx.synth_params(nx, ny, seed=21, flow=1, a0=100.0, a1=5.0, a2=0.002)
    a, b = pivx.synth_pair(sp_)
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=sp1, spacing_y1=sp1, npass1=2,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2)
    f = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    with pivx.Xcorr2dMulti(p) as x:
        x.set_mhe(True, 40, 0.91)
        x.prep_img_data(a, b)
        x.run_pair(f, f)
        r1 = [x.raw(2), x.filtered(2), x.filtered(3)]
    with pivx.Xcorr2dMulti(p) as y:
        y.prep_img_data(orc.mhe(a, 40, 0.91), orc.mhe(b, 40, 0.91))
        y.run_pair(f, f)
        r2 = [y.raw(2), y.filtered(2), y.filtered(3)]
    for A, B in zip(r1, r2):
        for u, w in zip(A, B):
            np.testing.assert_array_equal(u, w)


def test_mhe_rejects_untileable_settings(pivx):
    p = config1_params(pivx, nx=256, ny=128, npass2=0)
    with pivx.Xcorr2dMulti(p) as x:
        for ws, thr in ((0, 0.5), (129, 0.5), (16, 1.0)):
            with pytest.raises(pivx.PivError):
                x.set_mhe(True, ws, thr)
        with pytest.raises(pivx.PivError):
            x.mhe_enhance(np.zeros((128, 256), np.uint8))


def test_results_snapshot_survives_next_pair(pivx, tmp_path):
    """A result snapshot fetched after pair 1 still holds pair 1 while pair 2
    runs, and its files equal write_vec_data_general's bytes."""
    nx, ny = 512, 384
    p = config1_params(pivx, nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2,
                       ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2)
    f = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    a1, b1 = pivx.synth_pair(pivx.synth_params(nx, ny, seed=8, flow=1, a0=80.0, a1=4.0, a2=0.002))
    a2, b2 = pivx.synth_pair(pivx.synth_params(nx, ny, seed=9, flow=0, a0=2.0, a1=-1.0))
    with pivx.Xcorr2dMulti(p) as x, x.results() as r:
        x.prep_img_data(a1, b1)
        x.run_pair(f, f)
        direct = []
        for sel in range(5):
            path = str(tmp_path / f"d{sel}.dat")
            x.write_vec_data_general(path, sel, crlf=True)
            direct.append(open(path, "rb").read())
        r.fetch()
        x.prep_img_data_async(a2, b2)
        x.run_pair(f, f)
        r.wait()
        for sel in range(5):
            path = str(tmp_path / f"s{sel}.dat")
            r.write(path, sel, crlf=True)
            assert open(path, "rb").read() == direct[sel]
        x.synchronize()
