"""Pins the CPU oracle before it is trusted as the GPU checker.

Every oracle function is checked against something that does not share its
code: a direct DFT, brute-force enumeration on tiny inputs, closed forms of
the mathematics (Poisson summation makes the correlation of two sampled
Gaussians an exact Gaussian, on which the 3-point estimator is exact;
bilinear interpolation reproduces affine fields), numpy restatements, and the
reference's own fixtures (grid sizes of the committed result files)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [4, 8, 12, 20, 24, 28, 44, 52, 64, 96, 100, 124, 256, 1020, 1024])
def test_fft_matches_direct_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    ref = W @ x
    got = orc.fft1d(x, -1)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max() * np.log2(n)
    inv = orc.fft1d(x, +1)
    assert np.abs(inv - np.conj(W) @ x).max() <= 1e-12 * np.abs(ref).max() * np.log2(n)
    got32 = orc.fft1d(x, -1, prec=32)
    assert np.abs(got32 - ref).max() <= 1e-6 * np.abs(ref).max() * np.log2(n)


# ---------------------------------------------------------------- correlation plane
def brute_corr(f1, f2):
    """cc(s) = sum_x f1(x) f2(x + s) (circular), shifted so s = 0 sits at
    (ws_y/2, ws_x/2); an independent numpy restatement."""
    wsy, wsx = f1.shape
    cc = np.empty_like(f1, dtype=np.float64)
    for sy in range(wsy):
        for sx in range(wsx):
            cc[sy, sx] = np.sum(f1 * np.roll(f2, (-sy, -sx), axis=(0, 1)))
    return np.roll(np.abs(cc), (wsy // 2, wsx // 2), axis=(0, 1))


@pytest.mark.parametrize("shape", [(8, 8), (12, 20), (16, 16), (24, 8)])
def test_correlation_fft_equals_direct_and_brute_force(orc, shape):
    rng = np.random.default_rng(sum(shape))
    f1 = rng.integers(0, 256, shape).astype(float)
    f2 = rng.integers(0, 256, shape).astype(float)
    direct = orc.correlate_window(f1, f2, method=1)
    fft = orc.correlate_window(f1, f2, method=0)
    brute = brute_corr(f1, f2)
    np.testing.assert_array_equal(direct, brute)          # integer-exact in FP64
    assert np.abs(fft - brute).max() <= 1e-9 * brute.max()


def test_correlation_known_answers(orc):
    ws = (16, 32)  # (wsy, wsx)
    f1 = np.zeros(ws)
    f2 = np.zeros(ws)
    f1[5, 7] = 1
    f2[5 + 3, 7 - 4] = 1                                   # f2 = f1 displaced by (dx, dy) = (-4, +3)
    cc = orc.correlate_window(f1, f2, method=0)
    j, i = np.unravel_index(np.argmax(cc), ws)
    assert (i - ws[1] // 2, j - ws[0] // 2) == (-4, 3)
    assert abs(cc.max() - 1) < 1e-12 and np.sort(cc.ravel())[-2] < 1e-12
    c = orc.correlate_window(np.full(ws, 3.0), np.full(ws, 5.0), method=0)
    assert np.allclose(c, 15.0 * ws[0] * ws[1], rtol=1e-13)
    rng = np.random.default_rng(0)
    p = rng.integers(0, 256, ws).astype(float)
    q = np.roll(p, (2, 5), axis=(0, 1))                   # integer shift (dx, dy) = (5, 2)
    cc = orc.correlate_window(p, q, method=0)
    j, i = np.unravel_index(np.argmax(cc), ws)
    assert (i - ws[1] // 2, j - ws[0] // 2) == (5, 2)
    assert abs(cc.max() - np.sum(p * p)) < 1e-6


def _one_window_pass(orc, f1, f2, K=4, pct=0.4, pred=(0.0, 0.0)):
    ws = f1.shape[::-1]
    g = orc.geometry(ws[0], ws[1], 1, ws[0], ws[1], ws[0], ws[1], pct, pct)
    assert g["nx_vec"] == 1 and g["ny_vec"] == 1
    return orc.xcorr_pass(f1, f2u8=f2, geom=g, ws=ws, sp=ws, K=K,
                          pred=(np.array([pred[0]]), np.array([pred[1]])) if pred != (0.0, 0.0) else None)


def test_gaussian_subpixel_is_exact(orc):
    """Two sampled Gaussians (sigma 2) offset by d: by Poisson summation their
    discrete correlation is exactly Gaussian in s (aliasing ~ e^{-2 pi^2 sigma^2}),
    and the 3-point Gaussian estimator returns d.  Uses the deformed-frame
    input (real values) of a single 32x32 window."""
    ws = 32
    y, x = np.mgrid[0:ws, 0:ws].astype(float)
    for d in [(0.3, -0.45), (-2.7, 1.2), (4.49, 3.01)]:
        c = (15.5, 16.0)
        s = 2.0
        f1 = 200 * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2) / (2 * s * s))
        f2 = 200 * np.exp(-((x - c[0] - d[0]) ** 2 + (y - c[1] - d[1]) ** 2) / (2 * s * s))
        g = orc.geometry(ws, ws, 1, ws, ws, ws, ws, 0.4, 0.4)
        r = orc.xcorr_pass(np.round(f1).astype(np.uint8), f2def=f2, geom=g, ws=(ws, ws), sp=(ws, ws), K=1)
        # frame 1 must be u8 on this path, so use the real-valued pair through correlate_window instead
        cc = orc.correlate_window(f1, f2, method=0)
        j, i = np.unravel_index(np.argmax(cc), cc.shape)
        l0, lm, lp = np.log(cc[j, i]), np.log(cc[j, i - 1]), np.log(cc[j, i + 1])
        est = i + (lm - lp) / (lm + lp - 2 * l0) / 2 - ws / 2
        assert abs(est - d[0]) < 1e-9
        l0, lm, lp = np.log(cc[j, i]), np.log(cc[j - 1, i]), np.log(cc[j + 1, i])
        est = j + (lm - lp) / (lm + lp - 2 * l0) / 2 - ws / 2
        assert abs(est - d[1]) < 1e-9
        assert r["valid"][0, 0] == 1


@pytest.mark.parametrize("f", [(0.6, 1.0, 0.8), (0.999, 1.0, 0.9985), (120.0, 400.0, 390.0), (5.0, 5.0, 1.0)])
def test_fit_sensitivity_bounds_finite_differences(orc, f):
    """The ill-conditioned-fit bound equals the sum of the fit's partial
    derivatives (numpy central differences of the cu:334-340 formula) times
    the relative perturbation of each value, and bounds the change seen under
    every sign pattern of a small perturbation."""
    import itertools

    def fit(a, b, c):
        la, lb, lc = np.log(a), np.log(b), np.log(c)
        return (la - lc) / (la + lc - 2 * lb) / 2

    delta = 1e-7 * max(f)
    got = orc.fit_sensitivity(*f, delta)
    want = 0.0
    for i in range(3):
        h = 1e-6 * f[i]
        up, dn = list(f), list(f)
        up[i] += h
        dn[i] -= h
        want += abs((fit(*up) - fit(*dn)) / (2 * h)) * delta
    assert got == pytest.approx(want, rel=1e-4)
    base = fit(*f)
    worst = max(abs(fit(*(np.array(f) + delta * np.array(sg))) - base) for sg in itertools.product((-1, 1), repeat=3))
    assert worst <= got * (1 + 1e-3)
    assert worst >= got * 0.99  # first order: the corner pattern attains the bound
    # a flat peak is ill-conditioned, a sharp Gaussian peak is not
    assert orc.fit_sensitivity(0.9999, 1.0, 0.99985, 1e-6) > 1e-3 > orc.fit_sensitivity(0.4, 1.0, 0.5, 1e-6)


def _peaks_numpy(cc, bx, by, K, pred, lims):
    """Brute-force restatement of local maxima, ranking and the 3-point fit."""
    wsy, wsx = cc.shape
    cand = []
    for j in range(by, wsy - by):
        for i in range(bx, wsx - bx):
            v = cc[j, i]
            if v > 0 and v >= cc[j, i - 1] and v >= cc[j, i + 1] and v >= cc[j - 1, i] and v >= cc[j + 1, i]:
                cand.append((-v, j * wsx + i))
    cand.sort()
    out = []
    for k in range(K):
        if k >= len(cand):
            out.append((0.0, 0.0, 0, -1))
            continue
        v, p = -cand[k][0], cand[k][1]
        i, j = p % wsx, p // wsx
        l0 = np.log(v)
        a, b = np.log(cc[j, i - 1]), np.log(cc[j, i + 1])
        dx = i + (a - b) / (a + b - 2 * l0) / 2 - wsx / 2 + pred[0]
        a, b = np.log(cc[j - 1, i]), np.log(cc[j + 1, i])
        dy = j + (a - b) / (a + b - 2 * l0) / 2 - wsy / 2 + pred[1]
        ok = lims[0] <= dx <= lims[1] and lims[2] <= dy <= lims[3] and np.isfinite(dx) and np.isfinite(dy)
        out.append((dx if ok else 0.0, dy if ok else 0.0, int(ok), p))
    return out, len(cand)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_peaks_and_fit_match_brute_force(orc, pivx, seed):
    """xcorr_pass (gather + correlation + uncapped local maxima + top-K +
    fit) against numpy on a small frame, window by window."""
    sp = pivx.synth_params(96, 64, seed=seed, flow=0, a0=2.3, a1=-1.1)
    a, b = pivx.synth_pair(sp)
    g = orc.geometry(96, 64, 1, 16, 16, 8, 8, 0.4, 0.4)
    rng = np.random.default_rng(seed)
    nt = g["nx_vec"] * g["ny_vec"]
    px, py = rng.uniform(-1, 1, nt), rng.uniform(-1, 1, nt)
    lims = (-3.0, 3.5, -2.5, 100.0)
    r = orc.xcorr_pass(a, f2u8=b, geom=g, ws=(16, 16), sp=(8, 8), K=5, pred=(px, py), lims=lims)
    for w in range(nt):
        ix, iy = w % g["nx_vec"], w // g["nx_vec"]
        f1 = np.zeros((16, 16))
        f2 = np.zeros((16, 16))
        for rr in range(16):
            for cc_ in range(16):
                xi, yi = ix * 8 + cc_ - g["left_x"], iy * 8 + rr - g["left_y"]
                if 0 <= xi < 96 and 0 <= yi < 64:
                    f1[rr, cc_] = a[yi, xi]
                    f2[rr, cc_] = b[yi, xi]
        cc = brute_corr(f1, f2)
        exp, n = _peaks_numpy(cc, g["border_x"], g["border_y"], 5, (px[w], py[w]), lims)
        assert r["ncand"][w] == n
        for k, (dx, dy, ok, p) in enumerate(exp):
            assert r["pk_idx"][w, k] == p
            assert r["valid"][w, k] == ok
            assert abs(r["dx"][w, k] - dx) < 1e-9 and abs(r["dy"][w, k] - dy) < 1e-9


# ---------------------------------------------------------------- deformation, extension, upsampling
def _geom(orc, nx, ny, ws, sp):
    return orc.geometry(nx, ny, 1, ws, ws, sp, sp, 0.4, 0.4)


def test_extend_is_edge_replication(orc):
    rng = np.random.default_rng(0)
    fx, fy = rng.standard_normal((7, 9)), rng.standard_normal((7, 9))
    ex, ey = orc.extend(fx, fy, 3)
    np.testing.assert_array_equal(ex, np.pad(fx, 3, mode="edge"))
    np.testing.assert_array_equal(ey, np.pad(fy, 3, mode="edge"))


@pytest.mark.parametrize("prec", [64, 32])
def test_deform_identity_and_integer_shift(orc, prec):
    nx, ny = 200, 150
    img = np.random.default_rng(1).integers(0, 256, (ny, nx), dtype=np.uint8)
    g = _geom(orc, nx, ny, 32, 16)
    nvx, nvy = g["nx_vec"], g["ny_vec"]
    lx, ly = g["left_x"], g["left_y"]
    padded = np.zeros((g["ny_aug"] + 10, g["nx_aug"] + 10))
    padded[ly:ly + ny, lx:lx + nx] = img
    z = np.zeros((nvy + 6, nvx + 6))
    d = orc.deform(img, g, (32, 32), (16, 16), z, z, prec=prec)
    np.testing.assert_array_equal(d, padded[:g["ny_aug"], :g["nx_aug"]])
    # integer displacement (2, -3): out(x, y) = f2(x + 2, y - 3)
    d = orc.deform(img, g, (32, 32), (16, 16), z + 2, z - 3, prec=prec)
    big = np.zeros((g["ny_aug"] + 20, g["nx_aug"] + 20))
    big[ly + 10:ly + 10 + ny, lx + 10:lx + 10 + nx] = img
    exp = big[10 - 3:10 - 3 + g["ny_aug"], 10 + 2:10 + 2 + g["nx_aug"]]
    np.testing.assert_array_equal(d, exp)


def test_deform_affine_field_closed_form(orc):
    """An affine field sampled on the vector grid is reproduced exactly by the
    bilinear field interpolation; the image sample is then the bilinear
    interpolation of frame 2 at (x + u, y + v)."""
    nx, ny = 160, 128
    img = np.random.default_rng(2).integers(0, 256, (ny, nx), dtype=np.uint8).astype(float)
    g = _geom(orc, nx, ny, 32, 16)
    nvx, nvy = g["nx_vec"] + 6, g["ny_vec"] + 6
    xs = (32 // 2 - 3 * 16) + 16 * np.arange(nvx)       # ex-field sample positions, aug coordinates
    ys = (32 // 2 - 3 * 16) + 16 * np.arange(nvy)
    X, Y = np.meshgrid(xs, ys)
    ex = 0.7 + 0.01 * X - 0.004 * Y
    ey = -1.3 + 0.002 * X + 0.006 * Y
    d = orc.deform(img.astype(np.uint8), g, (32, 32), (16, 16), ex, ey, prec=64)
    ax, ay = np.meshgrid(np.arange(g["nx_aug"]), np.arange(g["ny_aug"]))
    sx = ax + 0.7 + 0.01 * ax - 0.004 * ay
    sy = ay - 1.3 + 0.002 * ax + 0.006 * ay
    ix, iy = np.floor(sx).astype(int), np.floor(sy).astype(int)
    fx, fy = sx - ix, sy - iy

    def px(x, y):
        xi, yi = x - g["left_x"], y - g["left_y"]
        ok = (xi >= 0) & (xi < nx) & (yi >= 0) & (yi < ny)
        return np.where(ok, img[np.clip(yi, 0, ny - 1), np.clip(xi, 0, nx - 1)], 0.0)

    exp = ((1 - fx) * (1 - fy) * px(ix, iy) + fx * (1 - fy) * px(ix + 1, iy) + (1 - fx) * fy * px(ix, iy + 1)
           + fx * fy * px(ix + 1, iy + 1))
    assert np.abs(d - exp).max() < 1e-9


def test_upsample_reproduces_affine_fields(orc):
    nx1, ny1 = 20, 14          # grid 1 at spacing 32
    k, j = np.meshgrid(np.arange(nx1 + 6), np.arange(ny1 + 6))
    X = 16 + 32 * (k - 3.0)
    Y = 16 + 32 * (j - 3.0)
    ex, ey = 1.5 + 0.01 * X, -0.5 + 0.02 * Y - 0.003 * X
    dx, dy = orc.upsample(ex, ey, (32, 32), (16, 16), (40, 28))
    k2, j2 = np.meshgrid(np.arange(40), np.arange(28))
    x2, y2 = (8 + 16 * k2).ravel(), (8 + 16 * j2).ravel()
    assert np.abs(dx - (1.5 + 0.01 * x2)).max() < 1e-12
    assert np.abs(dy - (-0.5 + 0.02 * y2 - 0.003 * x2)).max() < 1e-12
    c, _ = orc.upsample(np.full_like(ex, 2.25), ey, (32, 32), (16, 16), (40, 28))
    assert np.all(c == 2.25)


# ---------------------------------------------------------------- validation
def median_pass_numpy(sx, sy, sv, cx, cy, cv, K, nvx, nvy, wx, wy, minn, thr, eps=0.1):
    hx, hy = (wx - 1) // 2, (wy - 1) // 2
    ox, oy, ov = np.zeros(nvx * nvy), np.zeros(nvx * nvy), np.zeros(nvx * nvy, np.uint8)
    for t in range(nvx * nvy):
        vy, vx = divmod(t, nvx)
        X, Y = [], []
        for s in range(wx * wy):
            if s == (wx * wy - 1) // 2:
                continue
            qx, qy = vx - hx + s % wx, vy - hy + s // wx
            if 0 <= qx < nvx and 0 <= qy < nvy and sv[qy * nvx + qx]:
                X.append(sx[qy * nvx + qx])
                Y.append(sy[qy * nvx + qx])
        if len(X) < minn:
            continue
        mx, my = np.median(X), np.median(Y)
        rx, ry = np.median(np.abs(np.array(X) - mx)), np.median(np.abs(np.array(Y) - my))
        for k in range(K):
            o = t * K + k
            if not cv[o]:
                continue
            if abs(cx[o] - mx) / (rx + eps) < thr and abs(cy[o] - my) / (ry + eps) < thr:
                ox[t], oy[t], ov[t] = cx[o], cy[o], 1
                break
    return ox, oy, ov


@pytest.mark.parametrize("win", [(5, 5), (3, 3), (4, 6), (7, 3)])
def test_median_pass_matches_numpy(orc, win):
    rng = np.random.default_rng(win[0] * 7 + win[1])
    nvx, nvy, K = 23, 17, 3
    nt = nvx * nvy
    cx = 1.0 + 0.3 * rng.standard_normal(nt * K)
    cy = -2.0 + 0.3 * rng.standard_normal(nt * K)
    cx[rng.random(nt * K) < 0.1] += 6
    cv = (rng.random(nt * K) > 0.2).astype(np.uint8)
    sx, sy, sv = cx[::K].copy(), cy[::K].copy(), cv[::K].copy()
    gx, gy, gv, _ = orc.median_pass((sx, sy, sv), (cx, cy, cv), K, (nvx, nvy), win, 4, 2.0)
    ex, ey, ev = median_pass_numpy(sx, sy, sv, cx, cy, cv, K, nvx, nvy, win[0], win[1], 4, 2.0)
    np.testing.assert_array_equal(gv, ev)
    np.testing.assert_array_equal(gx, ex)
    np.testing.assert_array_equal(gy, ey)


def test_median_known_answers(orc):
    nvx, nvy, K = 7, 7, 3
    nt = nvx * nvy
    cx = np.tile([1.0, 9.0, 9.0], nt)
    cy = np.tile([2.0, 9.0, 9.0], nt)
    cv = np.ones(nt * K, np.uint8)
    c = 3 * nvx + 3
    cx[c * K:(c + 1) * K] = [20.0, -5.0, 1.05]   # outlier whose third candidate is consistent
    cy[c * K:(c + 1) * K] = [20.0, 7.0, 1.95]
    sx, sy, sv = cx[::K].copy(), cy[::K].copy(), cv[::K].copy()
    dx, dy, dv, _ = orc.median_pass((sx, sy, sv), (cx, cy, cv), K, (nvx, nvy), (5, 5), 5, 2.0)
    assert dv[c] == 1 and dx[c] == 1.05 and dy[c] == 1.95
    assert np.all(dv == 1) and np.all(dx[np.arange(nt) != c] == 1.0)
    # too few valid neighbours -> invalid and zero (cu:231-236)
    sv2 = np.zeros(nt, np.uint8)
    sv2[[0, 1, 2]] = 1
    dx, dy, dv, _ = orc.median_pass((sx, sy, sv2), (cx, cy, cv), K, (nvx, nvy), (5, 5), 5, 2.0)
    assert np.all(dv == 0) and np.all(dx == 0) and np.all(dy == 0)


def fill_numpy(dx, dy, dv, nvx, nvy):
    dx, dy, dv = dx.copy().reshape(nvy, nvx), dy.copy().reshape(nvy, nvx), dv.copy().reshape(nvy, nvx)
    sweeps = 0
    while True:
        sweeps += 1
        nx_, ny_, nf = dx.copy(), dy.copy(), np.zeros_like(dv)
        for j in range(nvy):
            for i in range(nvx):
                if dv[j, i]:
                    continue
                sl = (slice(max(0, j - 1), j + 2), slice(max(0, i - 1), i + 2))
                m = dv[sl] == 1
                if m.any():
                    nx_[j, i] = dx[sl][m].sum() / m.sum()
                    ny_[j, i] = dy[sl][m].sum() / m.sum()
                    nf[j, i] = 1
        if not nf.any():
            return dx.ravel(), dy.ravel(), dv.ravel(), sweeps
        dx, dy, dv = nx_, ny_, dv | nf


def test_fill_matches_numpy_and_known_answers(orc):
    rng = np.random.default_rng(4)
    nvx, nvy = 19, 13
    dx = rng.standard_normal(nvx * nvy)
    dy = rng.standard_normal(nvx * nvy)
    dv = (rng.random(nvx * nvy) > 0.6).astype(np.uint8)
    gx, gy, gv, gs = orc.fill(dx, dy, dv, (nvx, nvy))
    ex, ey, ev, es = fill_numpy(dx, dy, dv, nvx, nvy)
    assert gs == es
    np.testing.assert_array_equal(gv, ev)
    assert np.abs(gx - ex).max() < 1e-12 and np.abs(gy - ey).max() < 1e-12
    # isolated hole
    dv = np.ones(nvx * nvy, np.uint8)
    h = 5 * nvx + 7
    dv[h] = 0
    gx, _, gv, gs = orc.fill(dx, dy, dv, (nvx, nvy))
    nb = [h + a * nvx + b for a in (-1, 0, 1) for b in (-1, 0, 1) if (a, b) != (0, 0)]
    assert gs == 2 and gv[h] == 1 and abs(gx[h] - np.mean(dx[nb])) < 1e-12
    # nothing valid: unchanged after one sweep
    gx, _, gv, gs = orc.fill(np.zeros(12), np.zeros(12), np.zeros(12, np.uint8), (4, 3))
    assert gs == 1 and np.all(gv == 0)


def test_smooth_box_mean(orc):
    nvx, nvy = 15, 11
    k, j = np.meshgrid(np.arange(nvx), np.arange(nvy))
    sx = (0.5 + 0.1 * k - 0.2 * j).ravel()
    sy = (1.0 * k * 0 + 3.0).ravel()
    dx, dy = orc.smooth(sx, sy, (nvx, nvy), (3, 3))
    inner = ((k > 0) & (k < nvx - 1) & (j > 0) & (j < nvy - 1)).ravel()
    assert np.abs(dx - sx)[inner].max() < 1e-12 and np.all(np.abs(dy - 3.0) < 1e-12)
    rng = np.random.default_rng(1)
    r = rng.standard_normal(nvx * nvy)
    dx, _ = orc.smooth(r, r, (nvx, nvy), (3, 5))
    R = r.reshape(nvy, nvx)
    exp = np.empty_like(R)
    for jj in range(nvy):
        for ii in range(nvx):
            exp[jj, ii] = R[max(0, jj - 2):jj + 3, max(0, ii - 1):ii + 2].mean()
    assert np.abs(dx - exp.ravel()).max() < 1e-12


# ---------------------------------------------------------------- geometry
def test_geometry_matches_reference_result_files(orc):
    fixtures = json.load(open(os.path.join(GOLDEN, "result_fixtures.json")))
    g1 = orc.geometry(6600, 4400, 1, 64, 64, 32, 32, 0.4, 0.4)
    g2 = orc.geometry(6600, 4400, 1, 32, 32, 16, 16, 0.4, 0.4)
    for name, fx in fixtures.items():
        if "shiftvec" in name:
            continue
        g = g1 if ("Grid1" in name or "grid1" in name) else g2
        assert fx["header"][2].strip() == f'ZONE T="Frame 0", I={g["nx_vec"]}, J={g["ny_vec"]}'
        assert fx["n_lines"] == 3 + g["nx_vec"] * g["ny_vec"]
    assert (g1["nx_vec"], g1["ny_vec"], g2["nx_vec"], g2["ny_vec"]) == (206, 137, 412, 275)


def test_geometry_sheet(orc):
    """Config geometry: 512x512 @16 -> 32x32 vectors, pads 8/8, aug 528;
    6600x4400 64@64 -> 103x68 (no pad); 32@16 -> 412x275, pads (8,0)/(8,8)."""
    g = orc.geometry(512, 512, 1, 32, 32, 16, 16, 0.5, 0.5)
    assert (g["nx_vec"], g["ny_vec"], g["left_x"], g["right_x"], g["nx_aug"], g["border_x"], g["npeak_slots"]) == \
        (32, 32, 8, 8, 528, 1, 64)
    g = orc.geometry(6600, 4400, 1, 64, 64, 64, 64, 0.4, 0.4)
    assert (g["nx_vec"], g["ny_vec"], g["left_x"], g["right_x"], g["nx_aug"], g["ny_aug"], g["border_x"]) == \
        (103, 68, 0, 0, 6600, 4400, 6)
    g = orc.geometry(6600, 4400, 1, 32, 32, 16, 16, 0.4, 0.4)
    assert (g["left_x"], g["right_x"], g["left_y"], g["right_y"], g["nx_aug"], g["ny_aug"], g["border_x"]) == \
        (8, 0, 8, 8, 6608, 4416, 3)
    g = orc.geometry(1000, 1000, 3, 32, 32, 16, 16, 0.4, 0.4)  # 62*62 = 3844 vectors, 3 sections
    assert (g["neach"], g["nlast"], g["nsec"]) == (1281, 1, 4)


def test_run_pair_composes_the_stages(orc, pivx):
    """orc_run_pair (one grid, one pass) equals the stage functions composed
    by hand: prep, median passes, fill, smoothing."""
    sp = pivx.synth_params(160, 128, seed=9, flow=0, a0=2.0, a1=1.0)
    a, b = pivx.synth_pair(sp)
    ip = pivx.init_params(nx=160, ny=128, ws_x1=32, ws_y1=32, spacing_x1=16, spacing_y1=16, npass1=1, npass2=0,
                          ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16)
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    out = orc.run_pair(a, b, ip, fp, fp)
    g = orc.geometry(160, 128, 1, 32, 32, 16, 16, 0.4, 0.4)
    r = orc.xcorr_pass(a, f2u8=b, geom=g, ws=(32, 32), sp=(16, 16), K=4)
    np.testing.assert_array_equal(out["raw1"][0], r["dx"])
    nv = (g["nx_vec"], g["ny_vec"])
    cx, cy, cv = r["dx"].ravel(), r["dy"].ravel(), r["valid"].ravel()
    sx, sy, sv = r["dx"][:, 0], r["dy"][:, 0], r["valid"][:, 0]
    for _ in range(2):
        sx, sy, sv, _m = orc.median_pass((sx, sy, sv), (cx, cy, cv), 4, nv, (5, 5), 5, 2.0)
    sx, sy, sv, _s = orc.fill(sx, sy, sv, nv)
    sx, sy = orc.smooth(sx, sy, nv, (3, 3))
    np.testing.assert_array_equal(out["filt1"][0], sx)
    np.testing.assert_array_equal(out["filt1"][2], sv)


def test_run_pair_two_grid_composes_the_stages(orc, pivx):
    """orc_run_pair's two-grid branch (main:289-375) equals the stage
    functions composed by hand: grid-1 passes (pass 0 undeformed with a zero
    predictor, later passes deformed by the extended filtered field), the
    upsampled predictor, deformed grid-2 passes, and the median-only re-run
    of the last grid-2 candidates (no fill, no smoothing)."""
    nx, ny = 256, 192
    sp = pivx.synth_params(nx, ny, seed=19, flow=1, a0=200.0, a1=4.0, a2=0.002)
    a, b = pivx.synth_pair(sp)
    ip = pivx.init_params(nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2, ws_x2=32,
                          ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2)
    fp1 = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    fp2 = pivx.FilterParams(1, 4, 2.0, 3.0, 2)
    out = orc.run_pair(a, b, ip, fp1, fp2)
    geo = [orc.geometry(nx, ny, 1, 64, 64, 32, 32, 0.4, 0.4), orc.geometry(nx, ny, 1, 32, 32, 16, 16, 0.4, 0.4)]
    ws, sps = [(64, 64), (32, 32)], [(32, 32), (16, 16)]

    def filt(r, nv, f, fill=True, smooth=True):
        cand = (r["dx"].ravel(), r["dy"].ravel(), r["valid"].ravel())
        sx, sy, sv = r["dx"][:, 0], r["dy"][:, 0], r["valid"][:, 0]
        for _ in range(f.npass_median):
            sx, sy, sv, _m = orc.median_pass((sx, sy, sv), cand, 4, nv, (5, 5), f.min_nneigh, f.insert_thr)
        if fill:
            sx, sy, sv, _s = orc.fill(sx, sy, sv, nv)
        if smooth:
            for _ in range(f.npass_smooth):
                sx, sy = orc.smooth(sx, sy, nv, (3, 3))
        return sx, sy, sv

    nv = [(g["nx_vec"], g["ny_vec"]) for g in geo]
    fx = np.zeros(nv[0][0] * nv[0][1])
    fy = np.zeros_like(fx)
    for j in range(2):
        kw = dict(geom=geo[0], ws=ws[0], sp=sps[0], K=4, pred=(fx, fy))
        if j == 0:
            r = orc.xcorr_pass(a, f2u8=b, **kw)
        else:
            ex, ey = orc.extend(fx.reshape(nv[0][1], nv[0][0]), fy.reshape(nv[0][1], nv[0][0]))
            r = orc.xcorr_pass(a, f2def=orc.deform(b, geo[0], ws[0], sps[0], ex, ey), **kw)
        fx, fy, fv = filt(r, nv[0], fp1)
    np.testing.assert_array_equal(out["raw1"][0], r["dx"])
    np.testing.assert_array_equal(out["filt1"][0], fx)
    np.testing.assert_array_equal(out["filt1"][2], fv)
    ex, ey = orc.extend(fx.reshape(nv[0][1], nv[0][0]), fy.reshape(nv[0][1], nv[0][0]))
    gx, gy = orc.upsample(ex, ey, sps[0], sps[1], nv[1])
    for j in range(2):
        ex, ey = orc.extend(gx.reshape(nv[1][1], nv[1][0]), gy.reshape(nv[1][1], nv[1][0]))
        r = orc.xcorr_pass(a, f2def=orc.deform(b, geo[1], ws[1], sps[1], ex, ey), geom=geo[1], ws=ws[1], sp=sps[1],
                           K=4, pred=(gx, gy))
        gx, gy, gv = filt(r, nv[1], fp2)
    np.testing.assert_array_equal(out["raw2"][0], r["dx"])
    np.testing.assert_array_equal(out["raw2"][2], r["valid"])
    np.testing.assert_array_equal(out["filt2"][0], gx)
    np.testing.assert_array_equal(out["filt2"][1], gy)
    np.testing.assert_array_equal(out["filt2"][2], gv)
    mx, my, mv = filt(r, nv[1], fp2, fill=False, smooth=False)
    np.testing.assert_array_equal(out["mo2"][0], mx)
    np.testing.assert_array_equal(out["mo2"][2], mv)
    # the composed field follows the imposed flow (the branch is not vacuous)
    cx = 8 + 16 * np.arange(nv[1][0])
    cy = 8 + 16 * np.arange(nv[1][1])
    X, Y = np.meshgrid(cx, cy)
    u, v = pivx.synth_flow(sp, X.ravel(), Y.ravel())
    inner = ((X > 32) & (X < nx - 32) & (Y > 32) & (Y < ny - 32)).ravel()
    assert np.sqrt(np.mean((gx[inner] - u[inner]) ** 2 + (gy[inner] - v[inner]) ** 2)) < 0.2


# ---------------------------------------------------------------- MHE (MHE7_4_gpu.cu)
def _mhe_numpy(img, ws, thresh):
    """Vectorised restatement: per-tile np.bincount histograms, a zero ring,
    3x3 tile sums by slicing, np.cumsum over the levels, the enhanced-value
    formula and the tile lookup; pixels outside the tiled region are 0."""
    ny, nx = img.shape
    if nx % ws == 0 and ny % ws == 0:
        xs = ys = 0
    else:
        xs, ys = (nx - ws * (nx // ws)) // 2, (ny - ws * (ny // ws)) // 2
    ti, tj = ny // ws, nx // ws
    reg = img[ys:ys + ti * ws, xs:xs + tj * ws].astype(np.int64)
    tiles = reg.reshape(ti, ws, tj, ws).transpose(0, 2, 1, 3).reshape(ti, tj, ws * ws)
    hist = np.zeros((ti + 2, tj + 2, 256), np.int64)
    for a in range(ti):
        for b in range(tj):
            hist[a + 1, b + 1] = np.bincount(tiles[a, b], minlength=256)
    trans = sum(hist[1 + dy:1 + dy + ti, 1 + dx:1 + dx + tj] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    cum = np.cumsum(trans, axis=2)
    n = np.full((ti, tj), 6, np.int64)
    n[1:ti - 1, 1:tj - 1] = 9
    for a, b in ((0, 0), (ti - 1, 0), (0, tj - 1), (ti - 1, tj - 1)):
        n[a, b] = 4
    t = float(np.float32(ws * ws) * np.float32(thresh))
    ntrash = int(np.floor(abs(t) + 0.5)) * (1 if t >= 0 else -1)  # C roundf: half away from zero
    nn = ws * ws - ntrash
    c = np.maximum(cum - n[..., None] * ntrash, 0) * 255 // (n[..., None] * nn)
    lut = np.minimum(c, 255)
    lut[..., 0] = 0
    lut[..., 255] = 255
    out = np.zeros_like(img)
    ty, tx = np.meshgrid(np.arange(ti * ws) // ws, np.arange(tj * ws) // ws, indexing="ij")
    out[ys:ys + ti * ws, xs:xs + tj * ws] = lut[ty, tx, reg]
    return out


@pytest.mark.parametrize("nx,ny,ws,thr", [(96, 64, 16, 0.91), (101, 77, 12, 0.5), (64, 40, 8, 0.0),
                                          (50, 50, 50, 0.3), (120, 30, 20, 0.7)])
def test_mhe_matches_numpy(orc, nx, ny, ws, thr):
    rng = np.random.default_rng(nx + ny + ws)
    img = np.clip(rng.normal(40, 30, (ny, nx)), 0, 255).astype(np.uint8)
    img[rng.random((ny, nx)) < 0.05] = 255
    got = orc.mhe(img, ws, thr)
    np.testing.assert_array_equal(got, _mhe_numpy(img, ws, thr))


def test_mhe_known_answers(orc):
    # constant frame over a 2-D tiling: every level below the constant counts
    # nothing, the constant's cumulative count is exactly n*ws^2 -> 255
    img = np.full((48, 64), 77, np.uint8)
    assert (orc.mhe(img, 16, 0.91) == 255).all()
    # level 0 always maps to 0, level 255 to 255
    img[:8, :8] = 0
    img[-8:, -8:] = 255
    out = orc.mhe(img, 16, 0.91)
    assert (out[:8, :8] == 0).all() and (out[-8:, -8:] == 255).all()
    # one tile: its neighbour count is 4 (corner rule) although it has no
    # neighbours, so with threshold 0 level v of 64 distinct values maps to
    # floor(v*255/(4*64))
    img = (np.arange(64).reshape(8, 8) + 1).astype(np.uint8)
    out = orc.mhe(img, 8, 0.0)
    np.testing.assert_array_equal(out, (img.astype(int) * 255 // 256).astype(np.uint8))
    # a single row of three tiles: end tiles count 4 neighbours, the middle one
    # 6, while two and three tiles contribute -> a constant maps to 127 in all
    img = np.full((10, 30), 9, np.uint8)
    assert (orc.mhe(img, 10, 0.0) == 127).all()
    # sizes not divisible by the window: the central region is tiled, the
    # margins are 0
    img = np.full((45, 70), 33, np.uint8)
    out = orc.mhe(img, 20, 0.5)
    ys, xs = (45 - 40) // 2, (70 - 60) // 2
    assert (out[ys:ys + 40, xs:xs + 60] == 255).all()
    out[ys:ys + 40, xs:xs + 60] = 0
    assert (out == 0).all()
    with pytest.raises(ValueError):
        orc.mhe(img, 0, 0.5)
    with pytest.raises(ValueError):
        orc.mhe(img, 10, 1.0)
