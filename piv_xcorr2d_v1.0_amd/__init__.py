"""Python binding of the MI355X PIV engine (libpivx.so) over its C ABI.

The class :class:`Xcorr2dMulti` mirrors the reference's ``xcorr2d_multi``
(xcorr2d_V1.0/xcorr2d_multi_grid.cuh:12-208): same method names, same argument
meaning, same pass order; failures raise :class:`PivError` carrying the
engine's message instead of returning ``false``.  Arrays cross the boundary as
numpy arrays; nothing here computes on the CPU.

The directory name contains a dot, so load it by path::

    spec = importlib.util.spec_from_file_location("pivx", "<repo>/piv_xcorr2d_v1.0_amd/__init__.py")
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpivx.so")
CLI_PATH = os.path.join(_HERE, "piv_xcorr2d")

_lib = None


class PivError(RuntimeError):
    """A non-zero status from the C ABI (message from pivx_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[pivx error {code}] {msg}")
        self.code = code


class InitParams(C.Structure):
    """pivx_init_params: the 30 scalars of xcorr2d_multi::init (cu:716-718)."""

    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nsec_cc", "ws_x1", "ws_y1", "spacing_x1", "spacing_y1", "npass1",
        "ws_x2", "ws_y2", "spacing_x2", "spacing_y2", "npass2",
        "wx1_median", "wy1_median", "wx2_median", "wy2_median",
        "wx1_smooth", "wy1_smooth", "wx2_smooth", "wy2_smooth", "npeak_save")] + [
        (n, C.c_float) for n in (
            "max_dispx_percent1", "max_dispy_percent1", "max_dispx_percent2", "max_dispy_percent2",
            "dispx_lim1", "dispx_lim2", "dispy_lim1", "dispy_lim2")]


class Geometry(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "ws_x", "ws_y", "spacing_x", "spacing_y", "nx_vec", "ny_vec", "nt_vec", "border_x", "border_y",
        "npeak_slots", "left_x", "left_y", "right_x", "right_y", "nx_aug", "ny_aug", "nsec", "neach", "nlast")]


class FilterParams(C.Structure):
    _fields_ = [("npass_median", C.c_int), ("min_nneigh", C.c_int), ("residual_thr", C.c_float),
                ("insert_thr", C.c_float), ("npass_smooth", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("overflow_windows", C.c_longlong * 2), ("outliers", C.c_longlong * 2),
                ("fill_sweeps", C.c_int * 2), ("oob_taps", C.c_longlong * 2), ("hole_list_drops", C.c_longlong),
                ("holes_pending", C.c_int)]


class ParameterFile(C.Structure):
    _fields_ = [("prefix_f1", C.c_char * 1024), ("suffix_f1", C.c_char * 256), ("prefix_f2", C.c_char * 1024),
                ("suffix_f2", C.c_char * 256), ("dir_index", C.c_char * 1024), ("folder_result", C.c_char * 1024),
                ("device_id", C.c_int), ("ifmhe", C.c_int), ("mhe_ws", C.c_int), ("mhe_thresh", C.c_float),
                ("init", InitParams),
                ("npass1_median", C.c_int), ("min_nneigh_p1", C.c_int), ("residual_threshold_p1", C.c_float),
                ("insert_threshold_p1", C.c_float), ("npass1_smooth", C.c_int),
                ("npass2_median", C.c_int), ("min_nneigh_p2", C.c_int), ("residual_threshold_p2", C.c_float),
                ("insert_threshold_p2", C.c_float), ("npass2_smooth", C.c_int)]


class SynthParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("seed", C.c_ulonglong), ("density", C.c_float),
                ("sigma", C.c_float), ("imin", C.c_float), ("imax", C.c_float), ("background", C.c_float),
                ("noise_sigma", C.c_float), ("flow", C.c_int), ("a0", C.c_float), ("a1", C.c_float),
                ("a2", C.c_float)]


class CnnTrainParams(C.Structure):
    """pivx_cnn_train_params."""

    _fields_ = [("npairs", C.c_int), ("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float),
                ("shuffle_seed", C.c_ulonglong), ("synth", SynthParams), ("grid1", FilterParams)]


class CnnTrainReport(C.Structure):
    _fields_ = [("gen_seconds", C.c_double), ("train_seconds", C.c_double), ("steps", C.c_int),
                ("first_epoch_loss", C.c_float), ("last_epoch_loss", C.c_float)]


_P = C.c_void_p
_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

# name -> argtypes (restype int unless listed in _RESTYPE)
_SIGS = {
    "pivx_init": [C.POINTER(InitParams), C.c_int, C.POINTER(_P)],
    "pivx_freemem": [_P],
    "pivx_last_error": [_P],
    "pivx_get_geometry": [_P, C.c_int, C.POINTER(Geometry)],
    "pivx_compute_geometry": [C.POINTER(InitParams), C.c_int, C.POINTER(Geometry)],
    "pivx_prep_img_data": [_P, _P, _P, C.c_size_t],
    "pivx_prep_img_data_device": [_P, _P, _P, C.c_size_t],
    "pivx_set_mhe": [_P, C.c_int, C.c_int, C.c_float],
    "pivx_prep_img_data_async": [_P, _P, _P, C.c_size_t],
    "pivx_host_alloc": [C.c_size_t],
    "pivx_host_free": [_P],
    "pivx_results_create": [_P, C.POINTER(_P)],
    "pivx_results_fetch": [_P, _P],
    "pivx_results_wait": [_P],
    "pivx_results_write": [_P, C.c_char_p, C.c_int, C.c_int],
    "pivx_results_destroy": [_P],
    "pivx_mhe_enhance": [_P, _P, _P, C.c_size_t],
    "pivx_calc_cc_map_p1": [_P, C.c_int],
    "pivx_median_filter_smoothing_p1": [_P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int],
    "pivx_extend_filtered_vecfield_p1": [_P],
    "pivx_median_filter_smoothing_noalloc": [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                             _P, _P, _P, _P],
    "pivx_upsampling_vector_field": [_P],
    "pivx_extend_filtered_vecfield_p2": [_P],
    "pivx_calc_cc_map_p2": [_P],
    "pivx_median_filter_smoothing_p2": [_P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int],
    "pivx_run_pair": [_P, C.POINTER(FilterParams), C.POINTER(FilterParams)],
    "pivx_synchronize": [_P],
    "pivx_set_graph_mode": [_P, C.c_int],
    "pivx_set_fused_deform": [_P, C.c_int],
    "pivx_set_correlation_precision": [_P, C.c_int],
    "pivx_write_vec_data_general": [_P, C.c_char_p, C.c_int, C.c_int],
    "pivx_get_dims": [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pivx_get_grid_coords": [_P, C.c_int, _P, _P],
    "pivx_get_raw": [_P, C.c_int, _P, _P, _P],
    "pivx_get_filtered": [_P, C.c_int, _P, _P, _P],
    "pivx_get_peaks": [_P, C.c_int, _P, _P, _P],
    "pivx_get_deformed": [_P, C.c_int, _P],
    "pivx_get_stats": [_P, C.POINTER(Stats)],
    "pivx_set_filtered": [_P, C.c_int, _P, _P, _P],
    "pivx_set_raw": [_P, C.c_int, _P, _P, _P],
    "pivx_enable_timing": [_P, C.c_int],
    "pivx_get_stage_ms": [_P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "pivx_bandwidth_probe": [_P, C.c_size_t, C.c_int, C.POINTER(C.c_double)],
    "pivx_parse_parameter_file": [C.c_char_p, C.POINTER(ParameterFile)],
    "pivx_read_pair_index": [C.c_char_p, _P, C.c_int],
    "pivx_output_name": [C.POINTER(ParameterFile), C.c_int, C.c_int, C.c_char_p, C.c_size_t],
    "pivx_write_tecplot": [C.c_char_p, C.c_int, C.c_int, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int],
    "pivx_tiff_info": [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pivx_tiff_read8": [C.c_char_p, _P, C.c_size_t],
    "pivx_tiff_write8": [C.c_char_p, _P, C.c_int, C.c_int],
    "pivx_synth_pair": [C.POINTER(SynthParams), _P, _P],
    "pivx_synth_pair_device": [_P, C.POINTER(SynthParams)],
    "pivx_cnn_param_count": [],
    "pivx_cnn_create": [C.c_int, C.POINTER(_P)],
    "pivx_cnn_destroy": [_P],
    "pivx_cnn_init_params": [_P, C.c_ulonglong],
    "pivx_cnn_set_params": [_P, _P, C.c_int],
    "pivx_cnn_get_params": [_P, _P, C.c_int],
    "pivx_cnn_forward": [_P, _P, C.c_int, C.c_int, C.c_int, _P],
    "pivx_cnn_gradient": [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P],
    "pivx_cnn_adam_step": [_P, C.c_float],
    "pivx_cnn_train": [_P, _P, C.POINTER(CnnTrainParams), _P, C.POINTER(CnnTrainReport)],
    "pivx_run_pair_learned": [_P, _P, C.POINTER(FilterParams), C.POINTER(FilterParams), C.c_float],
    "pivx_get_learned_stats": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pivx_batch_create": [_P, C.c_int, C.POINTER(_P)],
    "pivx_batch_destroy": [_P],
    "pivx_batch_set_graph_mode": [_P, C.c_int],
    "pivx_batch_run": [_P, C.POINTER(FilterParams), C.POINTER(FilterParams)],
    "pivx_batch_synchronize": [_P],
    "pivx_batch_autotune": [_P, C.c_int],
    "pivx_batch_get_tuning": [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                              C.POINTER(C.c_longlong)],
    "pivx_synth_flow": [C.POINTER(SynthParams), _P, _P, C.c_int, _P, _P],
}
_RESTYPE = {"pivx_last_error": C.c_char_p, "pivx_host_alloc": C.c_void_p, "pivx_host_free": None,
            "pivx_results_destroy": None}


def lib():
    """Load libpivx.so from this package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {_HERE}` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def last_error(handle=None):
    return lib().pivx_last_error(handle).decode(errors="replace")


def _check(rc, handle=None):
    if rc != 0:
        raise PivError(rc, last_error(handle))


def init_params(**kw) -> InitParams:
    """InitParams with the sample Parameter.dat values as defaults (par:11-50)."""
    d = dict(nx=512, ny=512, nsec_cc=1, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=2,
             ws_x2=32, ws_y2=32, spacing_x2=16, spacing_y2=16, npass2=2,
             wx1_median=5, wy1_median=5, wx2_median=5, wy2_median=5,
             wx1_smooth=3, wy1_smooth=3, wx2_smooth=3, wy2_smooth=3, npeak_save=4,
             max_dispx_percent1=0.4, max_dispy_percent1=0.4, max_dispx_percent2=0.4, max_dispy_percent2=0.4,
             dispx_lim1=-100.0, dispx_lim2=100.0, dispy_lim1=-100.0, dispy_lim2=100.0)
    unknown = set(kw) - set(d)
    if unknown:
        raise KeyError(f"unknown init parameters {sorted(unknown)}")
    d.update(kw)
    return InitParams(**d)


class Xcorr2dMulti:
    """Mirror of class xcorr2d_multi over libpivx (one handle, one HIP stream)."""

    def __init__(self, params: InitParams, device_id: int = 0):
        self.p = params
        h = C.c_void_p()
        _check(lib().pivx_init(C.byref(params), device_id, C.byref(h)))
        self.h = h
        self.K = params.npeak_save

    # ------------------------------------------------------------ lifecycle
    def freemem(self):
        if getattr(self, "h", None):
            lib().pivx_freemem(self.h)
            self.h = None

    close = freemem

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.freemem()

    def __del__(self):
        try:
            self.freemem()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.h)

    def geometry(self, grid) -> Geometry:
        g = Geometry()
        self._c(lib().pivx_get_geometry(self.h, grid, C.byref(g)))
        return g

    # ------------------------------------------------------------ inputs
    def prep_img_data(self, img1: np.ndarray, img2: np.ndarray):
        a = np.ascontiguousarray(img1, dtype=np.uint8)
        b = np.ascontiguousarray(img2, dtype=np.uint8)
        if a.shape != (self.p.ny, self.p.nx) or b.shape != a.shape:
            raise ValueError(f"frames must be ({self.p.ny}, {self.p.nx}) uint8")
        self._c(lib().pivx_prep_img_data(self.h, _ptr(a), _ptr(b), self.p.nx))

    def prep_img_data_device(self, dptr1: int, dptr2: int, pitch: int):
        self._c(lib().pivx_prep_img_data_device(self.h, C.c_void_p(dptr1), C.c_void_p(dptr2), pitch))

    def prep_img_data_async(self, img1: np.ndarray, img2: np.ndarray):
        """Enqueue the uploads without waiting; keep img1/img2 alive and
        unchanged until synchronize() (pinned buffers overlap best)."""
        a = np.ascontiguousarray(img1, dtype=np.uint8)
        b = np.ascontiguousarray(img2, dtype=np.uint8)
        if a.shape != (self.p.ny, self.p.nx) or b.shape != a.shape:
            raise ValueError(f"frames must be ({self.p.ny}, {self.p.nx}) uint8")
        self._c(lib().pivx_prep_img_data_async(self.h, _ptr(a), _ptr(b), self.p.nx))
        self._keep = (a, b)

    def results(self):
        """A host result snapshot bound to this handle (see Results)."""
        return Results(self)

    # ------------------------------------------------------------ MHE (mhe_gpu, MHE7_4_gpu.cu:149-253)
    def set_mhe(self, enable: bool, ws: int = 40, thresh: float = 0.91):
        """Enhance every later uploaded frame on the GPU (main.cpp:267-283)."""
        self._c(lib().pivx_set_mhe(self.h, 1 if enable else 0, ws, thresh))

    def mhe_enhance(self, img: np.ndarray) -> np.ndarray:
        """One frame through the configured MHE; returns the enhanced frame."""
        a = np.ascontiguousarray(img, dtype=np.uint8)
        if a.shape != (self.p.ny, self.p.nx):
            raise ValueError(f"frame must be ({self.p.ny}, {self.p.nx}) uint8")
        out = np.empty_like(a)
        self._c(lib().pivx_mhe_enhance(self.h, _ptr(a), _ptr(out), self.p.nx))
        return out

    # ------------------------------------------------------------ passes (reference method names)
    def Calc_cc_map_p1(self, idx_pass: int):
        self._c(lib().pivx_calc_cc_map_p1(self.h, idx_pass))

    def Calc_cc_map_p2(self):
        self._c(lib().pivx_calc_cc_map_p2(self.h))

    def MedianFilter_Smoothing_p1(self, npass_median, min_nneigh, residual_thr, insert_thr, npass_smooth):
        self._c(lib().pivx_median_filter_smoothing_p1(self.h, npass_median, min_nneigh, residual_thr, insert_thr,
                                                       npass_smooth))

    def MedianFilter_Smoothing_p2(self, npass_median, min_nneigh, residual_thr, insert_thr, npass_smooth, iffill):
        self._c(lib().pivx_median_filter_smoothing_p2(self.h, npass_median, min_nneigh, residual_thr, insert_thr,
                                                       npass_smooth, iffill))

    def extend_filtered_vecfield_p1(self):
        self._c(lib().pivx_extend_filtered_vecfield_p1(self.h))

    def extend_filtered_vecfield_p2(self):
        self._c(lib().pivx_extend_filtered_vecfield_p2(self.h))

    def Upsampling_vector_field(self):
        self._c(lib().pivx_upsampling_vector_field(self.h))

    def run_pair(self, f1: FilterParams, f2: FilterParams):
        self._c(lib().pivx_run_pair(self.h, C.byref(f1), C.byref(f2)))

    def run_pair_learned(self, net: "Cnn", f1: FilterParams, f2: FilterParams, max_outlier_fraction: float = 0.05):
        """The pair sequence with the grid-1 passes replaced by the network
        (plain grid 1 as the fallback above the outlier fraction)."""
        self._c(lib().pivx_run_pair_learned(self.h, net.h, C.byref(f1), C.byref(f2), max_outlier_fraction))

    def learned_stats(self):
        """(outliers of the network field, 1 if the plain grid 1 ran instead)."""
        o, f = C.c_int(), C.c_int()
        self._c(lib().pivx_get_learned_stats(self.h, C.byref(o), C.byref(f)))
        return o.value, f.value

    def synth_pair_device(self, sp: SynthParams):
        """Render a synthetic pair on the GPU into this handle's frames."""
        self._c(lib().pivx_synth_pair_device(self.h, C.byref(sp)))

    def set_graph_mode(self, on=True):
        self._c(lib().pivx_set_graph_mode(self.h, 1 if on else 0))

    def set_correlation_precision(self, mode):
        """0 FP32 FFT; DFT-by-matrix on the matrix cores: 1 FP16, 2 BF16, 3 FP16 hi+lo, 4 BF16 hi+lo."""
        self._c(lib().pivx_set_correlation_precision(self.h, int(mode)))

    def set_fused_deform(self, mode=1):
        """0: never; 1: inside run_pair (default); 2: also in Calc_cc_map_p1."""
        self._c(lib().pivx_set_fused_deform(self.h, int(mode)))

    def synchronize(self):
        self._c(lib().pivx_synchronize(self.h))

    def write_vec_data_general(self, path: str, selection: int, crlf: bool = False):
        self._c(lib().pivx_write_vec_data_general(self.h, path.encode(), selection, 1 if crlf else 0))

    # ------------------------------------------------------------ outputs
    def dims(self, grid):
        g = self.geometry(grid)
        return g.nx_vec, g.ny_vec

    def grid_coords(self, grid):
        g = self.geometry(grid)
        x = np.empty(g.nt_vec, np.float32)
        y = np.empty(g.nt_vec, np.float32)
        self._c(lib().pivx_get_grid_coords(self.h, grid, _ptr(x), _ptr(y)))
        return x, y

    def raw(self, grid):
        n = self.geometry(grid).nt_vec
        dx = np.empty((n, self.K), np.float32)
        dy = np.empty((n, self.K), np.float32)
        v = np.empty((n, self.K), np.uint8)
        self._c(lib().pivx_get_raw(self.h, grid, _ptr(dx), _ptr(dy), _ptr(v)))
        return dx, dy, v

    def filtered(self, grid):
        """grid 1, 2, or 3 = the grid-2 median-only field of run_pair."""
        n = self.geometry(2 if grid == 3 else grid).nt_vec
        dx = np.empty(n, np.float32)
        dy = np.empty(n, np.float32)
        v = np.empty(n, np.uint8)
        self._c(lib().pivx_get_filtered(self.h, grid, _ptr(dx), _ptr(dy), _ptr(v)))
        return dx, dy, v

    def peaks(self, grid):
        n = self.geometry(grid).nt_vec
        val = np.empty((n, self.K), np.float32)
        idx = np.empty((n, self.K), np.int32)
        nc = np.empty(n, np.int32)
        self._c(lib().pivx_get_peaks(self.h, grid, _ptr(val), _ptr(idx), _ptr(nc)))
        return val, idx, nc

    def deformed(self, grid):
        g = self.geometry(grid)
        out = np.empty((g.ny_aug, g.nx_aug), np.float32)
        self._c(lib().pivx_get_deformed(self.h, grid, _ptr(out)))
        return out

    def stats(self) -> dict:
        s = Stats()
        self._c(lib().pivx_get_stats(self.h, C.byref(s)))
        return dict(overflow_windows=list(s.overflow_windows), outliers=list(s.outliers),
                    fill_sweeps=list(s.fill_sweeps), oob_taps=list(s.oob_taps),
                    hole_list_drops=int(s.hole_list_drops), holes_pending=int(s.holes_pending))

    def set_filtered(self, grid, dx, dy, valid=None):
        dx = np.ascontiguousarray(dx, np.float32)
        dy = np.ascontiguousarray(dy, np.float32)
        v = None if valid is None else np.ascontiguousarray(valid, np.uint8)
        self._c(lib().pivx_set_filtered(self.h, grid, _ptr(dx), _ptr(dy), _ptr(v)))

    def set_raw(self, grid, dx, dy, valid=None):
        dx = np.ascontiguousarray(dx, np.float32)
        dy = np.ascontiguousarray(dy, np.float32)
        v = None if valid is None else np.ascontiguousarray(valid, np.uint8)
        self._c(lib().pivx_set_raw(self.h, grid, _ptr(dx), _ptr(dy), _ptr(v)))

    def bandwidth_probe(self, nbytes=1 << 30, reps=20):
        """Streaming-copy GB/s (read + write) on this handle's GPU."""
        g = C.c_double()
        self._c(lib().pivx_bandwidth_probe(self.h, nbytes, reps, C.byref(g)))
        return g.value

    def enable_timing(self, on=True):
        self._c(lib().pivx_enable_timing(self.h, 1 if on else 0))

    def stage_ms(self, grid, stage):
        ms = C.c_double()
        n = C.c_int()
        self._c(lib().pivx_get_stage_ms(self.h, grid, stage, C.byref(ms), C.byref(n)))
        return ms.value, n.value


NOALLOC_MAX_VECTORS = 131072  # PIVX_NOALLOC_MAX_VECTORS


def median_filter_smoothing_noalloc(d_cand_x: int, d_cand_y: int, d_cand_valid: int, nx_vec: int, ny_vec: int,
                                    npeak_save: int, wx_median: int, wy_median: int, npass_median: int,
                                    min_nneigh: int, residual_thr: float, insert_thr: float, wx_smooth: int,
                                    wy_smooth: int, npass_smooth: int, iffill: int, d_out_x: int, d_out_y: int,
                                    d_out_valid: int, stream: int = 0):
    """pivx_median_filter_smoothing_noalloc on device addresses (ints):
    enqueues the validation stage on `stream` without allocating."""
    _check(lib().pivx_median_filter_smoothing_noalloc(
        C.c_void_p(d_cand_x), C.c_void_p(d_cand_y), C.c_void_p(d_cand_valid), nx_vec, ny_vec, npeak_save, wx_median,
        wy_median, npass_median, min_nneigh, residual_thr, insert_thr, wx_smooth, wy_smooth, npass_smooth, iffill,
        C.c_void_p(d_out_x), C.c_void_p(d_out_y), C.c_void_p(d_out_valid), C.c_void_p(stream)))


# ---------------------------------------------------------------- host utilities
def compute_geometry(params: InitParams, grid: int) -> Geometry:
    """The geometry pivx_init derives for a grid, without a device."""
    g = Geometry()
    _check(lib().pivx_compute_geometry(C.byref(params), grid, C.byref(g)))
    return g


def parse_parameter_file(path: str) -> ParameterFile:
    pf = ParameterFile()
    _check(lib().pivx_parse_parameter_file(path.encode(), C.byref(pf)))
    return pf


def read_pair_index(path: str):
    n = lib().pivx_read_pair_index(path.encode(), None, 0)
    if n < 0:
        raise PivError(4, last_error())
    buf = np.zeros(max(n, 1), np.int32)
    lib().pivx_read_pair_index(path.encode(), _ptr(buf), n)
    return [int(v) for v in buf[:n]]


def output_name(pf: ParameterFile, pair: int, which: int) -> str:
    buf = C.create_string_buffer(8192)
    _check(lib().pivx_output_name(C.byref(pf), pair, which, buf, len(buf)))
    return buf.value.decode()


def write_tecplot(path: str, nx_vec: int, ny_vec: int, gx, gy, dx, dy, raw: bool, crlf: bool = False):
    gx = np.ascontiguousarray(gx, np.float32)
    gy = np.ascontiguousarray(gy, np.float32)
    dx = np.ascontiguousarray(dx, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    K = dx.shape[1] if (raw and dx.ndim == 2) else 1
    _check(lib().pivx_write_tecplot(path.encode(), nx_vec, ny_vec, _ptr(gx), _ptr(gy), _ptr(dx), _ptr(dy), K,
                                    1 if raw else 0, 1 if crlf else 0))


def tiff_info(path: str):
    w, h, b = C.c_int(), C.c_int(), C.c_int()
    _check(lib().pivx_tiff_info(path.encode(), C.byref(w), C.byref(h), C.byref(b)))
    return w.value, h.value, b.value


def tiff_read8(path: str, nx: int, ny: int) -> np.ndarray:
    out = np.zeros((ny, nx), np.uint8)
    _check(lib().pivx_tiff_read8(path.encode(), _ptr(out), out.size))
    return out


def tiff_write8(path: str, img: np.ndarray):
    a = np.ascontiguousarray(img, np.uint8)
    _check(lib().pivx_tiff_write8(path.encode(), _ptr(a), a.shape[1], a.shape[0]))


def synth_params(nx, ny, seed=1, density=0.05, sigma=0.6, imin=120.0, imax=255.0, background=10.0,
                 noise_sigma=3.0, flow=0, a0=0.0, a1=0.0, a2=0.0) -> SynthParams:
    return SynthParams(nx=nx, ny=ny, seed=seed, density=density, sigma=sigma, imin=imin, imax=imax,
                       background=background, noise_sigma=noise_sigma, flow=flow, a0=a0, a1=a1, a2=a2)


def synth_pair(sp: SynthParams):
    a = np.empty((sp.ny, sp.nx), np.uint8)
    b = np.empty((sp.ny, sp.nx), np.uint8)
    _check(lib().pivx_synth_pair(C.byref(sp), _ptr(a), _ptr(b)))
    return a, b


def synth_flow(sp: SynthParams, x, y):
    x = np.ascontiguousarray(x, np.float32).ravel()
    y = np.ascontiguousarray(y, np.float32).ravel()
    u = np.empty_like(x)
    v = np.empty_like(x)
    _check(lib().pivx_synth_flow(C.byref(sp), _ptr(x), _ptr(y), x.size, _ptr(u), _ptr(v)))
    return u, v


class Cnn:
    """The learned grid-1 predictor (pivx_cnn_*): parameters live on the GPU."""

    def __init__(self, device_id: int = 0, seed: int = None):
        h = C.c_void_p()
        _check(lib().pivx_cnn_create(device_id, C.byref(h)))
        self.h = h
        self.n = lib().pivx_cnn_param_count()
        if seed is not None:
            _check(lib().pivx_cnn_init_params(self.h, seed))

    def close(self):
        if getattr(self, "h", None):
            lib().pivx_cnn_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(self.n, np.float32)
        _check(lib().pivx_cnn_get_params(self.h, _ptr(p), self.n))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().pivx_cnn_set_params(self.h, _ptr(p), p.size))

    def forward(self, x0):
        x0 = np.ascontiguousarray(x0, np.float32)
        b, h, w, _ = x0.shape
        out = np.empty((b, 2, (h // 8) * (w // 8)), np.float32)
        _check(lib().pivx_cnn_forward(self.h, _ptr(x0), b, h, w, _ptr(out)))
        return out

    def gradient(self, x0, labels):
        x0 = np.ascontiguousarray(x0, np.float32)
        labels = np.ascontiguousarray(labels, np.float32)
        b, h, w, _ = x0.shape
        loss = C.c_float()
        g = np.empty(self.n, np.float32)
        _check(lib().pivx_cnn_gradient(self.h, _ptr(x0), _ptr(labels), b, h, w, C.byref(loss), _ptr(g)))
        return loss.value, g

    def adam_step(self, lr):
        _check(lib().pivx_cnn_adam_step(self.h, lr))

    def train(self, x: "Xcorr2dMulti", tp: CnnTrainParams):
        losses = np.zeros(tp.epochs, np.float32)
        rep = CnnTrainReport()
        x._c(lib().pivx_cnn_train(x.h, self.h, C.byref(tp), _ptr(losses), C.byref(rep)))
        return losses, rep


class Batch:
    """Batched passes over many handles (pivx_batch_*): one persistent
    correlation launch per pass over all pairs."""

    def __init__(self, handles):
        self.handles = list(handles)
        arr = (C.c_void_p * len(self.handles))(*[x.h.value for x in self.handles])
        h = C.c_void_p()
        _check(lib().pivx_batch_create(arr, len(self.handles), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pivx_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph_mode(self, on=True):
        _check(lib().pivx_batch_set_graph_mode(self.h, 1 if on else 0))

    def run(self, f1: FilterParams, f2: FilterParams):
        _check(lib().pivx_batch_run(self.h, C.byref(f1), C.byref(f2)))

    def synchronize(self):
        _check(lib().pivx_batch_synchronize(self.h))

    def autotune(self, reps=3):
        _check(lib().pivx_batch_autotune(self.h, reps))

    def tuning(self, grid):
        w, c, ms, n = C.c_int(), C.c_int(), C.c_double(), C.c_longlong()
        _check(lib().pivx_batch_get_tuning(self.h, grid, C.byref(w), C.byref(c), C.byref(ms), C.byref(n)))
        return dict(waves_per_cu=w.value, items_per_claim=c.value, ms=ms.value, items=n.value)


class Results:
    """Page-locked snapshot of the fields the five output files need:
    fetch() enqueues the copies, wait() blocks until they land, write()
    formats one file from the snapshot (selection 0..4 as
    write_vec_data_general) without touching the GPU."""

    def __init__(self, x: "Xcorr2dMulti"):
        self.x = x
        r = _P()
        x._c(lib().pivx_results_create(x.h, C.byref(r)))
        self.r = r

    def fetch(self):
        self.x._c(lib().pivx_results_fetch(self.x.h, self.r))

    def wait(self):
        _check(lib().pivx_results_wait(self.r), None)

    def write(self, path: str, selection: int, crlf: bool = False):
        _check(lib().pivx_results_write(self.r, path.encode(), selection, 1 if crlf else 0), None)

    def close(self):
        if self.r:
            lib().pivx_results_destroy(self.r)
            self.r = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

