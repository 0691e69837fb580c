"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference stages.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_D = C.c_void_p


class OrcPass(C.Structure):
    _fields_ = [("f1", C.c_void_p), ("f2u8", C.c_void_p), ("f2def", C.c_void_p),
                ("nx", C.c_int), ("ny", C.c_int), ("nx_aug", C.c_int), ("ny_aug", C.c_int),
                ("wsx", C.c_int), ("wsy", C.c_int), ("spx", C.c_int), ("spy", C.c_int),
                ("left_x", C.c_int), ("left_y", C.c_int), ("nx_vec", C.c_int), ("ny_vec", C.c_int),
                ("bx", C.c_int), ("by", C.c_int), ("K", C.c_int),
                ("lx1", C.c_double), ("lx2", C.c_double), ("ly1", C.c_double), ("ly2", C.c_double),
                ("pred_x", C.c_void_p), ("pred_y", C.c_void_p), ("method", C.c_int), ("tie_rel", C.c_double),
                ("threads", C.c_int), ("win_begin", C.c_int), ("win_end", C.c_int), ("fit_tol", C.c_double),
                ("fit_rel", C.c_double)]


class OrcRun(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nsec", "ws_x1", "ws_y1", "sp_x1", "sp_y1", "npass1", "ws_x2", "ws_y2", "sp_x2", "sp_y2",
        "npass2", "wx1_median", "wy1_median", "wx2_median", "wy2_median", "wx1_smooth", "wy1_smooth",
        "wx2_smooth", "wy2_smooth", "K")] + [
        (n, C.c_double) for n in ("pct_x1", "pct_y1", "pct_x2", "pct_y2", "lx1", "lx2", "ly1", "ly2")] + [
        ("npass1_median", C.c_int), ("min_nneigh1", C.c_int), ("thr1", C.c_double), ("npass1_smooth", C.c_int),
        ("npass2_median", C.c_int), ("min_nneigh2", C.c_int), ("thr2", C.c_double), ("npass2_smooth", C.c_int),
        ("threads", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built; run `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        L.orc_geometry.argtypes = [C.c_int] * 7 + [C.c_float, C.c_float, _D]
        L.orc_correlate_window.argtypes = [C.c_int, _D, _D, C.c_int, C.c_int, C.c_int, _D]
        L.orc_fft1d.argtypes = [C.c_int, _D, C.c_int, C.c_int]
        L.orc_xcorr_pass.argtypes = [C.c_int, C.POINTER(OrcPass), _D, _D, _D, _D, _D, _D, _D]
        L.orc_xcorr_stage_seconds.argtypes = [C.POINTER(OrcPass), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]
        L.orc_deform.argtypes = [C.c_int, _D] + [C.c_int] * 6 + [_D, _D, C.c_int, C.c_int] + [C.c_double] * 4 + [_D]
        L.orc_extend.argtypes = [_D, _D, C.c_int, C.c_int, C.c_int, _D, _D]
        L.orc_upsample.argtypes = [C.c_int, _D, _D, C.c_int, C.c_int] + [C.c_double] * 4 + [C.c_int] * 4 + [_D, _D]
        L.orc_median_pass.argtypes = ([C.c_int, _D, _D, _D, C.c_int, _D, _D, _D] + [C.c_int] * 6 +
                                      [C.c_double, C.c_double, _D, _D, _D, _D])
        L.orc_fill.argtypes = [C.c_int, _D, _D, _D, C.c_int, C.c_int, _D]
        L.orc_smooth.argtypes = [C.c_int, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int, _D, _D]
        L.orc_run_pair.argtypes = [C.c_int, C.POINTER(OrcRun), _D, _D] + [_D] * 15
        L.orc_mhe.argtypes = [_D, _D, C.c_int, C.c_int, C.c_int, C.c_float]
        L.orc_cnn_forward.argtypes = [_D, _D, C.c_int, C.c_int, C.c_int, C.c_int, _D]
        L.orc_cnn_grad.argtypes = [_D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int, _D, _D]
        L.orc_fit_sensitivity.argtypes = [C.c_double] * 4
        L.orc_fit_sensitivity.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def geometry(nx, ny, nsec, wsx, wsy, spx, spy, pctx, pcty):
    out = np.zeros(14, np.int32)
    rc = lib().orc_geometry(nx, ny, nsec, wsx, wsy, spx, spy, pctx, pcty, _p(out))
    assert rc == 0
    keys = ("nx_vec", "ny_vec", "border_x", "border_y", "npeak_slots", "left_x", "left_y", "right_x", "right_y",
            "nx_aug", "ny_aug", "nsec", "neach", "nlast")
    return dict(zip(keys, (int(v) for v in out)))


def correlate_window(f1w, f2w, method=0, prec=64):
    f1w = _f64(f1w)
    f2w = _f64(f2w)
    wsy, wsx = f1w.shape
    cc = np.empty((wsy, wsx), np.float64)
    lib().orc_correlate_window(prec, _p(f1w), _p(f2w), wsx, wsy, method, _p(cc))
    return cc


def fft1d(x, sign=-1, prec=64):
    a = np.ascontiguousarray(np.asarray(x, np.complex128))
    buf = np.empty(2 * a.size, np.float64)
    buf[0::2] = a.real
    buf[1::2] = a.imag
    lib().orc_fft1d(prec, _p(buf), a.size, sign)
    return buf[0::2] + 1j * buf[1::2]


def xcorr_pass(f1, f2u8=None, f2def=None, *, geom, ws, sp, K, lims=(-100, 100, -100, 100), pred=None, method=0,
               tie_rel=0.0, prec=64, threads=1, win_range=None, fit_tol=1e-3, fit_rel=None):
    """One correlation pass of a grid.  geom: dict from geometry(); ws, sp: (x, y).
    tie [nt]: 1 where values within tie_rel * peak can change the candidates,
    2 where values within fit_rel * peak (default: one FP32 rounding per FFT
    stage, 2^-23 * log2(N), when tie_rel is set) can instead move a valid
    sub-pixel fit by more than fit_tol px."""
    if fit_rel is None:
        fit_rel = 2.0 ** -23 * np.log2(ws[0] * ws[1]) if tie_rel > 0 else 0.0
    f1 = np.ascontiguousarray(f1, np.uint8)
    ny, nx = f1.shape
    f2u8 = None if f2u8 is None else np.ascontiguousarray(f2u8, np.uint8)
    f2def = _f64(f2def)
    nt = geom["nx_vec"] * geom["ny_vec"]
    px = py = None
    if pred is not None:
        px, py = _f64(pred[0]), _f64(pred[1])
    P = OrcPass(f1=_p(f1), f2u8=_p(f2u8), f2def=_p(f2def), nx=nx, ny=ny, nx_aug=geom["nx_aug"],
                ny_aug=geom["ny_aug"], wsx=ws[0], wsy=ws[1], spx=sp[0], spy=sp[1], left_x=geom["left_x"],
                left_y=geom["left_y"], nx_vec=geom["nx_vec"], ny_vec=geom["ny_vec"], bx=geom["border_x"],
                by=geom["border_y"], K=K, lx1=lims[0], lx2=lims[1], ly1=lims[2], ly2=lims[3], pred_x=_p(px),
                pred_y=_p(py), method=method, tie_rel=tie_rel, threads=threads,
                win_begin=0 if win_range is None else win_range[0], win_end=0 if win_range is None else win_range[1],
                fit_tol=fit_tol, fit_rel=fit_rel)
    dx = np.zeros((nt, K))
    dy = np.zeros((nt, K))
    v = np.zeros((nt, K), np.uint8)
    pv = np.zeros((nt, K))
    pi = np.full((nt, K), -1, np.int32)
    nc = np.zeros(nt, np.int32)
    tie = np.zeros(nt, np.uint8)
    rc = lib().orc_xcorr_pass(prec, C.byref(P), _p(dx), _p(dy), _p(v), _p(pv), _p(pi), _p(nc), _p(tie))
    assert rc == 0
    return dict(dx=dx, dy=dy, valid=v, pk_val=pv, pk_idx=pi, ncand=nc, tie=tie)


def cnn_param_count():
    return lib().orc_cnn_param_count()


def cnn_forward(params, x0, bf16=False):
    """Network output [batch][2][h/8 * w/8] for input x0 [batch][h][w][2]."""
    params = _f64(params)
    x0 = _f64(x0)
    b, h, w, _ = x0.shape
    out = np.empty((b, 2, (h // 8) * (w // 8)))
    assert lib().orc_cnn_forward(_p(params), _p(x0), b, h, w, 1 if bf16 else 0, _p(out)) == 0
    return out


def cnn_grad(params, x0, labels, bf16=False):
    """(loss, dL/dparams) of the mean squared error against labels [batch][2][h/8 * w/8]."""
    params = _f64(params)
    x0 = _f64(x0)
    labels = _f64(labels)
    b, h, w, _ = x0.shape
    loss = C.c_double()
    g = np.empty(params.size)
    assert lib().orc_cnn_grad(_p(params), _p(x0), _p(labels), b, h, w, 1 if bf16 else 0, C.byref(loss), _p(g)) == 0
    return loss.value, g


def fit_sensitivity(f0, f1, f2, delta):
    """Bound on the change of one axis's 3-point fit when its values move by delta."""
    return lib().orc_fit_sensitivity(f0, f1, f2, delta)


def xcorr_stage_seconds(f1, f2u8=None, f2def=None, *, geom, ws, sp, K, pred=None):
    """(gather, transforms, peaks) seconds of one single-threaded FP32 pass."""
    f1 = np.ascontiguousarray(f1, np.uint8)
    ny, nx = f1.shape
    f2u8 = None if f2u8 is None else np.ascontiguousarray(f2u8, np.uint8)
    f2def = _f64(f2def)
    px = py = None
    if pred is not None:
        px, py = _f64(pred[0]), _f64(pred[1])
    P = OrcPass(f1=_p(f1), f2u8=_p(f2u8), f2def=_p(f2def), nx=nx, ny=ny, nx_aug=geom["nx_aug"],
                ny_aug=geom["ny_aug"], wsx=ws[0], wsy=ws[1], spx=sp[0], spy=sp[1], left_x=geom["left_x"],
                left_y=geom["left_y"], nx_vec=geom["nx_vec"], ny_vec=geom["ny_vec"], bx=geom["border_x"],
                by=geom["border_y"], K=K, lx1=-100, lx2=100, ly1=-100, ly2=100, pred_x=_p(px), pred_y=_p(py),
                method=0, tie_rel=0.0, threads=1, win_begin=0, win_end=0, fit_tol=0.0, fit_rel=0.0)
    t = [C.c_double() for _ in range(3)]
    assert lib().orc_xcorr_stage_seconds(C.byref(P), *[C.byref(v) for v in t]) == 0
    return tuple(v.value for v in t)


def deform(f2, geom, ws, sp, ex, ey, prec=64):
    f2 = np.ascontiguousarray(f2, np.uint8)
    ny, nx = f2.shape
    ex = _f64(ex)
    ey = _f64(ey)
    nye, nxe = ex.shape
    out = np.empty((geom["ny_aug"], geom["nx_aug"]), np.float64)
    lib().orc_deform(prec, _p(f2), nx, ny, geom["left_x"], geom["left_y"], geom["nx_aug"], geom["ny_aug"], _p(ex),
                     _p(ey), nxe, nye, float(ws[0] // 2 - 3 * sp[0]), float(ws[1] // 2 - 3 * sp[1]), float(sp[0]),
                     float(sp[1]), _p(out))
    return out


def extend(fx, fy, b=3):
    fx = _f64(fx)
    fy = _f64(fy)
    ny, nx = fx.shape
    ex = np.empty((ny + 2 * b, nx + 2 * b))
    ey = np.empty_like(ex)
    lib().orc_extend(_p(fx), _p(fy), nx, ny, b, _p(ex), _p(ey))
    return ex, ey


def upsample(ex, ey, sp1, sp2, n2, prec=64):
    ex = _f64(ex)
    ey = _f64(ey)
    nye, nxe = ex.shape
    nx2, ny2 = n2
    dx = np.empty(nx2 * ny2)
    dy = np.empty(nx2 * ny2)
    lib().orc_upsample(prec, _p(ex), _p(ey), nxe, nye, float(sp1[0] // 2 - 3 * sp1[0]), float(sp1[1] // 2 - 3 * sp1[1]),
                       float(sp1[0]), float(sp1[1]), sp2[0], sp2[1], nx2, ny2, _p(dx), _p(dy))
    return dx, dy


def median_pass(src, cand, K, nvec, win, min_nneigh, thr, eps=0.1, sstride=1, prec=64):
    """src = (sx, sy, sv) neighbourhood field; cand = (cx, cy, cv) [nt*K]."""
    sx, sy = _f64(src[0]), _f64(src[1])
    sv = np.ascontiguousarray(src[2], np.uint8)
    cx, cy = _f64(cand[0]), _f64(cand[1])
    cv = np.ascontiguousarray(cand[2], np.uint8)
    nt = nvec[0] * nvec[1]
    dx = np.empty(nt)
    dy = np.empty(nt)
    dv = np.empty(nt, np.uint8)
    mg = np.empty(nt)
    lib().orc_median_pass(prec, _p(sx), _p(sy), _p(sv), sstride, _p(cx), _p(cy), _p(cv), K, nvec[0], nvec[1], win[0],
                          win[1], min_nneigh, thr, eps, _p(dx), _p(dy), _p(dv), _p(mg))
    return dx, dy, dv, mg


def fill(dx, dy, dv, nvec, prec=64):
    dx = np.array(dx, np.float64)
    dy = np.array(dy, np.float64)
    dv = np.array(dv, np.uint8)
    sw = C.c_int()
    lib().orc_fill(prec, _p(dx), _p(dy), _p(dv), nvec[0], nvec[1], C.byref(sw))
    return dx, dy, dv, sw.value


def smooth(sx, sy, nvec, win, prec=64):
    sx = _f64(sx)
    sy = _f64(sy)
    dx = np.empty(nvec[0] * nvec[1])
    dy = np.empty_like(dx)
    lib().orc_smooth(prec, _p(sx), _p(sy), nvec[0], nvec[1], win[0], win[1], _p(dx), _p(dy))
    return dx, dy


def mhe(img, ws, thresh):
    """MHE-enhanced copy of one u8 frame (MHE7_4_gpu.cu)."""
    a = np.ascontiguousarray(img, np.uint8)
    out = np.empty_like(a)
    if lib().orc_mhe(_p(a), _p(out), a.shape[1], a.shape[0], ws, thresh) != 0:
        raise ValueError("MHE window or threshold out of range")
    return out


def run_pair(f1, f2, ip, fp1, fp2, prec=64, threads=1):
    """Whole pair.  ip: object with the init-parameter attribute names of the
    engine (nx, ws_x1, ...); fp1/fp2: objects with npass_median, min_nneigh,
    insert_thr, npass_smooth."""
    R = OrcRun(nx=ip.nx, ny=ip.ny, nsec=ip.nsec_cc, ws_x1=ip.ws_x1, ws_y1=ip.ws_y1, sp_x1=ip.spacing_x1,
               sp_y1=ip.spacing_y1, npass1=ip.npass1, ws_x2=ip.ws_x2, ws_y2=ip.ws_y2, sp_x2=ip.spacing_x2,
               sp_y2=ip.spacing_y2, npass2=ip.npass2, wx1_median=ip.wx1_median, wy1_median=ip.wy1_median,
               wx2_median=ip.wx2_median, wy2_median=ip.wy2_median, wx1_smooth=ip.wx1_smooth,
               wy1_smooth=ip.wy1_smooth, wx2_smooth=ip.wx2_smooth, wy2_smooth=ip.wy2_smooth, K=ip.npeak_save,
               pct_x1=ip.max_dispx_percent1, pct_y1=ip.max_dispy_percent1, pct_x2=ip.max_dispx_percent2,
               pct_y2=ip.max_dispy_percent2, lx1=ip.dispx_lim1, lx2=ip.dispx_lim2, ly1=ip.dispy_lim1,
               ly2=ip.dispy_lim2, npass1_median=fp1.npass_median, min_nneigh1=fp1.min_nneigh, thr1=fp1.insert_thr,
               npass1_smooth=fp1.npass_smooth, npass2_median=fp2.npass_median, min_nneigh2=fp2.min_nneigh,
               thr2=fp2.insert_thr, npass2_smooth=fp2.npass_smooth, threads=threads)
    g1 = geometry(ip.nx, ip.ny, ip.nsec_cc, ip.ws_x1, ip.ws_y1, ip.spacing_x1, ip.spacing_y1,
                  ip.max_dispx_percent1, ip.max_dispy_percent1)
    g2 = geometry(ip.nx, ip.ny, ip.nsec_cc, ip.ws_x2, ip.ws_y2, ip.spacing_x2, ip.spacing_y2,
                  ip.max_dispx_percent2, ip.max_dispy_percent2)
    K = ip.npeak_save
    n1 = g1["nx_vec"] * g1["ny_vec"]
    n2 = g2["nx_vec"] * g2["ny_vec"]
    out = dict(raw1=(np.zeros((n1, K)), np.zeros((n1, K)), np.zeros((n1, K), np.uint8)),
               filt1=(np.zeros(n1), np.zeros(n1), np.zeros(n1, np.uint8)),
               raw2=(np.zeros((n2, K)), np.zeros((n2, K)), np.zeros((n2, K), np.uint8)),
               filt2=(np.zeros(n2), np.zeros(n2), np.zeros(n2, np.uint8)),
               mo2=(np.zeros(n2), np.zeros(n2), np.zeros(n2, np.uint8)))
    f1 = np.ascontiguousarray(f1, np.uint8)
    f2 = np.ascontiguousarray(f2, np.uint8)
    args = []
    for k in ("raw1", "filt1", "raw2", "filt2", "mo2"):
        args += [_p(a) for a in out[k]]
    rc = lib().orc_run_pair(prec, C.byref(R), _p(f1), _p(f2), *args)
    assert rc == 0
    return out
