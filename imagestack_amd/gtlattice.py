"""ctypes binding of libgtlattice.so (include/gtlattice.h).

Device-pointer entry points take torch tensors on a ROCm device (torch is used
only for device memory and streams); the *_host entry points take numpy
arrays.  There is no CPU fallback: if the library or a GPU is missing, the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgtlattice.so")

GT_STORE_F32, GT_STORE_BF16, GT_STORE_F16 = 0, 1, 2
STORAGE = {"f32": GT_STORE_F32, "bf16": GT_STORE_BF16, "f16": GT_STORE_F16}
ERRORS = {1: "GT_EINVAL", 2: "GT_ENOMEM", 3: "GT_ERANGE", 4: "GT_EDEVICE", 5: "GT_ESTATE", 6: "GT_EUNSUPPORTED",
          7: "GT_ECANCELED", 8: "GT_ETIMEDOUT"}
GT_EINVAL, GT_ENOMEM, GT_ESTATE, GT_EUNSUPPORTED, GT_ECANCELED, GT_ETIMEDOUT = 1, 2, 5, 6, 7, 8

EXPORTS = [
    "gt_ctx_create", "gt_ctx_destroy", "gt_ctx_set_stream", "gt_ctx_set_profiling", "gt_last_error",
    "gt_gauss_transform", "gt_gauss_transform_host", "gt_joint_bilateral", "gt_joint_bilateral_host",
    "gt_joint_bilateral_noalloc", "gt_build_lattice", "gt_lattice_export", "gt_joint_bilateral_batch",
    "gt_autotune_joint_bilateral", "gt_autotune_joint_bilateral_batch",
    "gt_server_create", "gt_server_destroy", "gt_server_submit", "gt_server_wait", "gt_server_cancel",
    "gt_server_replay", "gt_debug_fail_allocations", "gt_gauss_transform_grid", "gt_gauss_transform_grid_host", "gt_joint_bilateral_grid",
    "gt_joint_bilateral_grid_host", "gt_patch_project", "gt_nlmeans", "gt_nlmeans_host",
    "gt_tmp_info", "gt_tmp_read", "gt_tmp_write", "gt_tmp_load", "gt_tmp_save",
    "gt_net_server_create", "gt_net_server_port", "gt_net_server_stats", "gt_net_server_destroy",
]

# .tmp type codes (src/FileTMP.cpp:14) and their numpy types
TMP_TYPES = {"float32": 0, "float64": 1, "uint8": 2, "int8": 3, "uint16": 4, "int16": 5, "uint32": 6,
             "int32": 7, "uint64": 8, "int64": 9}
GT_NET_ECHO, GT_NET_BILATERAL = 0, 1


class GTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Tuning(C.Structure):
    _fields_ = [("load_factor", C.c_float), ("waves_per_cu", C.c_int), ("dedup", C.c_int), ("tile_w", C.c_int),
                ("graph", C.c_int), ("neighbours", C.c_int), ("parts", C.c_int)]


class TuneReport(C.Structure):
    _fields_ = [("trials", C.c_int), ("best_ms", C.c_float), ("default_ms", C.c_float)]


def tuning_dict(t: Tuning) -> dict:
    return {k: getattr(t, k) for k, _ in Tuning._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int), ("max_batch", C.c_int), ("batch_pixels", C.c_int), ("hbm_budget", C.c_size_t),
                ("warm_side", C.c_int)]


class Request(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("C", C.c_int), ("im", C.c_void_p), ("row_pitch", C.c_longlong),
                ("plane_pitch", C.c_longlong), ("out", C.c_void_p), ("sigma_s", C.c_float), ("sigma_r", C.c_float),
                ("storage", C.c_int), ("deadline_ms", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("batch", C.c_int), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
                ("footprint", C.c_size_t)]


class NetConfig(C.Structure):
    _fields_ = [("port", C.c_int), ("reply_port", C.c_int), ("op", C.c_int), ("sigma_s", C.c_float),
                ("sigma_r", C.c_float), ("storage", C.c_int), ("max_pixels", C.c_longlong),
                ("serving", ServerConfig), ("max_bytes", C.c_longlong), ("max_inflight", C.c_longlong),
                ("bind_addr", C.c_char_p)]


class NetStats(C.Structure):
    _fields_ = [("received", C.c_longlong), ("replied", C.c_longlong), ("failed", C.c_longlong),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("refused", C.c_longlong),
                ("inflight_bytes", C.c_longlong)]


class BilateralImage(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("im", C.c_void_p), ("ref", C.c_void_p), ("out", C.c_void_p),
                ("sigma_s", C.c_float), ("sigma_r", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("M", C.c_int), ("capacity", C.c_int), ("retries", C.c_int), ("sort_passes", C.c_int),
                ("ms_splat", C.c_float), ("ms_sort", C.c_float), ("ms_nbr", C.c_float), ("ms_blur", C.c_float),
                ("ms_slice", C.c_float), ("ms_total", C.c_float), ("probes", C.c_ulonglong),
                ("inserts", C.c_ulonglong), ("atomics", C.c_ulonglong)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


def lib():
    """Load libgtlattice.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        L.gt_last_error.restype = C.c_char_p
        L.gt_ctx_create.argtypes = [C.POINTER(_vp), C.c_int]
        L.gt_ctx_destroy.argtypes = [_vp]
        L.gt_ctx_destroy.restype = None
        L.gt_ctx_set_stream.argtypes = [_vp, _vp]
        L.gt_ctx_set_profiling.argtypes = [_vp, C.c_int]
        sig = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp, _fp, _vp, C.c_int,
               C.POINTER(Tuning), C.POINTER(Stats)]
        L.gt_gauss_transform.argtypes = sig
        L.gt_gauss_transform_host.argtypes = sig
        jsig = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_float, C.c_float,
                C.c_float, C.c_float, _vp, C.c_int, C.POINTER(Tuning), C.POINTER(Stats)]
        L.gt_joint_bilateral.argtypes = jsig
        L.gt_joint_bilateral_host.argtypes = jsig
        L.gt_joint_bilateral_noalloc.argtypes = jsig
        L.gt_build_lattice.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _fp, C.c_int,
                                       C.POINTER(Tuning), C.POINTER(C.c_int)]
        L.gt_lattice_export.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.gt_joint_bilateral_batch.argtypes = [_vp, C.c_int, C.POINTER(BilateralImage), C.c_int, C.c_int, C.c_int,
                                               C.POINTER(Tuning), C.POINTER(Stats)]
        L.gt_autotune_joint_bilateral.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp,
                                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                                  C.POINTER(Tuning), C.POINTER(TuneReport)]
        L.gt_autotune_joint_bilateral_batch.argtypes = [_vp, C.c_int, C.POINTER(BilateralImage), C.c_int, C.c_int,
                                                        C.c_int, C.c_int, C.POINTER(Tuning), C.POINTER(TuneReport)]
        gsig = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp, _fp, _vp, C.POINTER(Stats)]
        L.gt_gauss_transform_grid.argtypes = gsig
        L.gt_gauss_transform_grid_host.argtypes = gsig
        jgsig = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_float, C.c_float, C.c_float,
                 C.c_float, _vp, C.POINTER(Stats)]
        L.gt_joint_bilateral_grid.argtypes = jgsig
        L.gt_joint_bilateral_grid_host.argtypes = jgsig
        L.gt_patch_project.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                                       C.c_int, _vp, _vp]
        nsig = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_float,
                C.c_float, _vp, C.c_int, C.POINTER(Tuning), C.POINTER(Stats)]
        L.gt_nlmeans.argtypes = nsig
        L.gt_nlmeans_host.argtypes = nsig
        L.gt_tmp_info.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.gt_tmp_read.argtypes = [C.c_char_p, _vp, C.c_size_t, C.POINTER(C.c_int)]
        L.gt_tmp_write.argtypes = [C.c_char_p, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.gt_tmp_load.argtypes = [_vp, C.c_char_p, _vp, C.c_size_t, C.POINTER(C.c_int)]
        L.gt_tmp_save.argtypes = [_vp, C.c_char_p, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.gt_net_server_create.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(NetConfig)]
        L.gt_net_server_port.argtypes = [_vp]
        L.gt_net_server_stats.argtypes = [_vp, C.POINTER(NetStats)]
        L.gt_net_server_destroy.argtypes = [_vp]
        L.gt_net_server_destroy.restype = None
        L.gt_server_create.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(ServerConfig)]
        L.gt_server_destroy.argtypes = [_vp]
        L.gt_server_destroy.restype = None
        L.gt_server_submit.argtypes = [_vp, C.POINTER(Request), C.POINTER(C.c_longlong)]
        L.gt_server_wait.argtypes = [_vp, C.c_longlong, C.POINTER(RequestResult)]
        L.gt_server_cancel.argtypes = [_vp, C.c_longlong]
        L.gt_debug_fail_allocations.argtypes = [C.c_int, C.c_int]
        L.gt_server_replay.argtypes = [_vp, C.c_int, C.POINTER(Request), C.POINTER(C.c_double),
                                       C.POINTER(RequestResult)]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise GTError(rc, lib().gt_last_error().decode())


def fail_allocations(count: int, after: int = 0) -> None:
    """gt_debug_fail_allocations: of the library's next device or pinned
    allocations, `after` succeed and then `count` fail (0 turns it off).
    For failure-path tests."""
    _check(lib().gt_debug_fail_allocations(after, count))


def _dptr(t) -> int:
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous() or t.dtype.itemsize != 4:
        raise ValueError("expected a contiguous 32-bit device tensor")
    return t.data_ptr()


def _tuning(tune) -> "Tuning | None":
    if tune is None:
        return None
    if isinstance(tune, Tuning):
        return tune
    t = Tuning()
    for k, v in dict(tune).items():
        setattr(t, k, v)
    return t


class Context:
    """One gt_ctx: device buffers reused across calls, one HIP stream."""

    def __init__(self, device: int = 0, profiling: int = 0):
        self._h = _vp()
        _check(lib().gt_ctx_create(C.byref(self._h), device))
        if profiling:
            _check(lib().gt_ctx_set_profiling(self._h, int(profiling)))

    def close(self) -> None:
        if self._h:
            lib().gt_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, level: int) -> None:
        """0 off, 1 per-stage HIP-event times, 2 also probe/atomic counters."""
        _check(lib().gt_ctx_set_profiling(self._h, int(level)))

    def set_stream(self, stream_ptr: int | None) -> None:
        _check(lib().gt_ctx_set_stream(self._h, stream_ptr))

    # --- device-pointer API (torch tensors) -----------------------------
    def gauss_transform(self, splat, values, inv_sigma, slice_pos=None, out=None, row_w: int = 0,
                        storage: str = "f32", tune=None, stats: bool = False):
        """splat: (d, N) device f32, values: (vd, N); returns (vd, Ns)."""
        import torch
        d, n = splat.shape[0], splat[0].numel()
        vd = values.shape[0]
        ns = n if slice_pos is None else slice_pos[0].numel()
        if out is None:
            out = torch.empty((vd, ns), dtype=torch.float32, device=splat.device)
        inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
        st = Stats()
        _check(lib().gt_gauss_transform(self._h, n, row_w, d, vd, _dptr(splat), _dptr(values), ns,
                                        _dptr(slice_pos), inv.ctypes.data_as(_fp), _dptr(out),
                                        STORAGE[storage], _tuning(tune), C.byref(st)))
        return (out, st.as_dict()) if stats else out

    def joint_bilateral(self, im, ref, fw, fh, ff, color_sigma, out=None, storage: str = "f32",
                        tune=None, stats: bool = False):
        """im: (C, F, H, W) device f32, ref: (Cr, F, H, W); returns (C, F, H, W)."""
        import torch
        Cc, F, H, W = im.shape
        if out is None:
            out = torch.empty_like(im)
        st = Stats()
        _check(lib().gt_joint_bilateral(self._h, W, H, F, Cc, _dptr(im), ref.shape[0], _dptr(ref), fw, fh,
                                        ff, color_sigma, _dptr(out), STORAGE[storage], _tuning(tune),
                                        C.byref(st)))
        return (out, st.as_dict()) if stats else out

    def joint_bilateral_noalloc(self, im, ref, fw, fh, ff, color_sigma, out, storage: str = "f32",
                                tune=None, stats: bool = False):
        """gt_joint_bilateral_noalloc: as joint_bilateral, into a given `out`
        (C, F, H, W) device f32; the call allocates no device memory."""
        Cc, F, H, W = im.shape
        st = Stats()
        _check(lib().gt_joint_bilateral_noalloc(self._h, W, H, F, Cc, _dptr(im), ref.shape[0], _dptr(ref), fw, fh,
                                                ff, color_sigma, _dptr(out), STORAGE[storage], _tuning(tune),
                                                C.byref(st)))
        return (out, st.as_dict()) if stats else out

    def joint_bilateral_batch(self, items, storage: str = "f32", tune=None, stats: bool = False):
        """items: list of (im (C,1,H,W), ref (Cr,1,H,W), out, sigma_s, sigma_r)
        device tensors; filters every image in one batched lattice."""
        arr, C_, Cr = self._batch_array(items)
        st = Stats()
        _check(lib().gt_joint_bilateral_batch(self._h, len(items), arr, C_, Cr, STORAGE[storage], _tuning(tune),
                                              C.byref(st)))
        return st.as_dict() if stats else None

    def gauss_transform_grid(self, splat, values, inv_sigma, slice_pos=None, out=None, stats: bool = False):
        """GaussTransform::apply(GRID): splat (d, N) device f32, values (vd, N);
        returns (vd, Ns) unnormalised."""
        import torch
        d, n = splat.shape[0], splat[0].numel()
        vd = values.shape[0]
        ns = n if slice_pos is None else slice_pos[0].numel()
        if out is None:
            out = torch.empty((vd, ns), dtype=torch.float32, device=splat.device)
        inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
        st = Stats()
        _check(lib().gt_gauss_transform_grid(self._h, n, d, vd, _dptr(splat), _dptr(values), ns, _dptr(slice_pos),
                                             inv.ctypes.data_as(_fp), _dptr(out), C.byref(st)))
        return (out, st.as_dict()) if stats else out

    def joint_bilateral_grid(self, im, ref, fw, fh, ff, color_sigma, out=None, stats: bool = False):
        """JointBilateral::apply(GRID): im (C, F, H, W) device f32, ref (Cr, F, H, W)."""
        import torch
        Cc, F, H, W = im.shape
        if out is None:
            out = torch.empty_like(im)
        st = Stats()
        _check(lib().gt_joint_bilateral_grid(self._h, W, H, F, Cc, _dptr(im), ref.shape[0], _dptr(ref), fw, fh,
                                             ff, color_sigma, _dptr(out), C.byref(st)))
        return (out, st.as_dict()) if stats else out

    def joint_bilateral_grid_host(self, im, ref, fw, fh, ff, color_sigma):
        im = np.ascontiguousarray(im, dtype=np.float32)
        ref = np.ascontiguousarray(ref, dtype=np.float32)
        Cc, F, H, W = im.shape
        out = np.zeros_like(im)
        st = Stats()
        _check(lib().gt_joint_bilateral_grid_host(self._h, W, H, F, Cc, im.ctypes.data, ref.shape[0],
                                                  ref.ctypes.data, fw, fh, ff, color_sigma, out.ctypes.data,
                                                  C.byref(st)))
        return out

    def patch_project(self, im, filters, out=None):
        """Convolve::apply(im, filters, Zero, Inner): im (C, F, H, W) device f32,
        filters (K*C, ft, fh, fw) device f32; returns (K, F, H, W)."""
        import torch
        Cc, F, H, W = im.shape
        KC, ft, fh, fw = filters.shape
        if KC % Cc:
            raise ValueError("filter planes must be a multiple of the image channels")
        K = KC // Cc
        if out is None:
            out = torch.empty((K, F, H, W), dtype=torch.float32, device=im.device)
        _check(lib().gt_patch_project(self._h, W, H, F, Cc, _dptr(im), fw, fh, ft, K, _dptr(filters), _dptr(out)))
        return out

    def nlmeans(self, im, filters, spatial_sigma, patch_sigma, out=None, storage: str = "f32", tune=None,
                stats: bool = False):
        """NLMeans::apply after the PCA: im (C, F, H, W) device f32, filters
        (K*C, ft, fh, fw) device f32 (PatchPCA's output); returns (C, F, H, W)."""
        import torch
        Cc, F, H, W = im.shape
        KC, ft, fh, fw = filters.shape
        if KC % Cc:
            raise ValueError("filter planes must be a multiple of the image channels")
        if out is None:
            out = torch.empty_like(im)
        st = Stats()
        _check(lib().gt_nlmeans(self._h, W, H, F, Cc, _dptr(im), fw, fh, ft, KC // Cc, _dptr(filters), spatial_sigma,
                                patch_sigma, _dptr(out), STORAGE[storage], _tuning(tune), C.byref(st)))
        return (out, st.as_dict()) if stats else out

    def nlmeans_host(self, im, filters, spatial_sigma, patch_sigma, storage: str = "f32"):
        """nlmeans on numpy arrays (host memory)."""
        im = np.ascontiguousarray(im, dtype=np.float32)
        filt = np.ascontiguousarray(filters, dtype=np.float32)
        Cc, F, H, W = im.shape
        KC, ft, fh, fw = filt.shape
        if KC % Cc:
            raise ValueError("filter planes must be a multiple of the image channels")
        out = np.zeros_like(im)
        st = Stats()
        _check(lib().gt_nlmeans_host(self._h, W, H, F, Cc, im.ctypes.data, fw, fh, ft, KC // Cc, filt.ctypes.data,
                                     spatial_sigma, patch_sigma, out.ctypes.data, STORAGE[storage], None,
                                     C.byref(st)))
        return out

    def tmp_load(self, path: str):
        """Loads a .tmp file into a new (C, F, H, W) float32 device tensor
        (samples cross PCIe in the file's type, converted on the GPU)."""
        import torch
        W, H, F, Cc = tmp_info(path)[0]
        out = torch.empty((Cc, F, H, W), dtype=torch.float32, device="cuda")
        dims = (C.c_int * 4)()
        _check(lib().gt_tmp_load(self._h, os.fsencode(path), _dptr(out), out.numel(), dims))
        return out

    def tmp_save(self, path: str, im, type_name: str = "float32") -> None:
        """Saves a (C, F, H, W) float32 device tensor as a .tmp file of the
        given sample type (converted on the GPU)."""
        Cc, F, H, W = im.shape
        _check(lib().gt_tmp_save(self._h, os.fsencode(path), _dptr(im.contiguous()), W, H, F, Cc,
                                 TMP_TYPES[type_name]))

    @staticmethod
    def _batch_array(items):
        arr = (BilateralImage * len(items))()
        C_ = Cr = None
        for k, (im, ref, out, ss, sr) in enumerate(items):
            Cc, F, H, W = im.shape
            assert F == 1 and out.shape == im.shape
            C_, Cr = Cc, ref.shape[0]
            arr[k] = BilateralImage(W, H, _dptr(im), _dptr(ref), _dptr(out), ss, sr)
        return arr, C_ or 1, Cr or 1

    def autotune_joint_bilateral(self, im, ref, fw, fh, ff, color_sigma, storage: str = "f32", reps: int = 3):
        """Search tile width x load factor x waves per CU on these inputs;
        returns (tuning dict, report dict).  `im` is only read."""
        Cc, F, H, W = im.shape
        best, rep = Tuning(), TuneReport()
        _check(lib().gt_autotune_joint_bilateral(self._h, W, H, F, Cc, _dptr(im), ref.shape[0], _dptr(ref), fw, fh,
                                                 ff, color_sigma, STORAGE[storage], reps, C.byref(best),
                                                 C.byref(rep)))
        return tuning_dict(best), {k: getattr(rep, k) for k, _ in TuneReport._fields_}

    def autotune_joint_bilateral_batch(self, items, storage: str = "f32", reps: int = 3):
        arr, C_, Cr = self._batch_array(items)
        best, rep = Tuning(), TuneReport()
        _check(lib().gt_autotune_joint_bilateral_batch(self._h, len(items), arr, C_, Cr, STORAGE[storage], reps,
                                                       C.byref(best), C.byref(rep)))
        return tuning_dict(best), {k: getattr(rep, k) for k, _ in TuneReport._fields_}

    def build_lattice(self, splat, values, inv_sigma, row_w: int = 0, storage: str = "f32", tune=None,
                      parts=("keys", "nbr", "splat", "blur")):
        """Splat + blur and export: dict(keys (M,d) int16, nbr (d+1,M,2) int32,
        splat (M,vd) f32, blur (M,vd) f32); `parts` names the arrays copied
        out (the others are None)."""
        d, n = splat.shape[0], splat[0].numel()
        vd = values.shape[0]
        inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
        M = C.c_int(0)
        _check(lib().gt_build_lattice(self._h, n, row_w, d, vd, _dptr(splat), _dptr(values),
                                      inv.ctypes.data_as(_fp), STORAGE[storage], _tuning(tune), C.byref(M)))
        M = M.value
        keys = np.zeros((M, d), np.int16) if "keys" in parts else None
        nbr = np.zeros((d + 1, M, 2), np.int32) if "nbr" in parts else None
        sv = np.zeros((M, vd), np.float32) if "splat" in parts else None
        bv = np.zeros((M, vd), np.float32) if "blur" in parts else None

        def ptr(a):
            return None if a is None else a.ctypes.data

        _check(lib().gt_lattice_export(self._h, ptr(keys), ptr(nbr), ptr(sv), ptr(bv)))
        return {"M": M, "keys": keys, "nbr": nbr, "splat": sv, "blur": bv}

    # --- host-pointer API (numpy) -----------------------------------------
    def gauss_transform_host(self, splat, values, inv_sigma, slice_pos=None, row_w: int = 0,
                             storage: str = "f32", tune=None):
        sp = np.ascontiguousarray(splat, dtype=np.float32)
        va = np.ascontiguousarray(values, dtype=np.float32)
        d = sp.shape[0]
        n = sp[0].size
        vd = va.shape[0]
        sl = None if slice_pos is None else np.ascontiguousarray(slice_pos, dtype=np.float32)
        ns = n if sl is None else sl[0].size
        out = np.zeros((vd, ns), np.float32)
        inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
        st = Stats()
        _check(lib().gt_gauss_transform_host(self._h, n, row_w, d, vd, sp.ctypes.data, va.ctypes.data, ns,
                                             None if sl is None else sl.ctypes.data, inv.ctypes.data_as(_fp),
                                             out.ctypes.data, STORAGE[storage], _tuning(tune), C.byref(st)))
        return out

    def joint_bilateral_host(self, im, ref, fw, fh, ff, color_sigma, storage: str = "f32", tune=None):
        im = np.ascontiguousarray(im, dtype=np.float32)
        ref = np.ascontiguousarray(ref, dtype=np.float32)
        Cc, F, H, W = im.shape
        out = np.zeros_like(im)
        st = Stats()
        _check(lib().gt_joint_bilateral_host(self._h, W, H, F, Cc, im.ctypes.data, ref.shape[0],
                                             ref.ctypes.data, fw, fh, ff, color_sigma, out.ctypes.data,
                                             STORAGE[storage], _tuning(tune), C.byref(st)))
        return out


def _host_ptr(a) -> int:
    """Address of a float32 host buffer: a numpy array or a CPU torch tensor
    (pinned or not)."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float32:
            raise ValueError("expected float32")
        return a.ctypes.data
    if getattr(a, "is_cuda", False) or a.dtype.itemsize != 4:
        raise ValueError("expected a float32 host tensor")
    return a.data_ptr()


def make_request(im, out, sigma_s: float, sigma_r: float, storage: str = "f32", row_pitch: int | None = None,
                 plane_pitch: int | None = None, W: int | None = None, H: int | None = None,
                 deadline_ms: float = 0.0) -> Request:
    """A bilateral request on host memory.  `im` is (C, H, W) dense float32, or
    (with W, H and the pitches given) a view into a larger planar buffer whose
    first pixel `im` points at; `out` is dense (C, H, W)."""
    Cc = out.shape[0]
    Hh = H if H is not None else out.shape[-2]
    Ww = W if W is not None else out.shape[-1]
    rp = row_pitch if row_pitch is not None else Ww
    pp = plane_pitch if plane_pitch is not None else Ww * Hh
    addr = im if isinstance(im, int) else _host_ptr(im)
    return Request(Ww, Hh, Cc, addr, rp, pp, _host_ptr(out), sigma_s, sigma_r, STORAGE[storage], deadline_ms)


class Server:
    """gt_server: a request loop with admission by HBM footprint, continuous
    batching and `slots` concurrent workers."""

    def __init__(self, device: int = 0, slots: int = 4, max_batch: int = 16, batch_pixels: int = 1 << 22,
                 hbm_budget: int = 0, warm_side: int = 0):
        self._h = _vp()
        cfg = ServerConfig(slots, max_batch, batch_pixels, hbm_budget, warm_side)
        _check(lib().gt_server_create(C.byref(self._h), device, C.byref(cfg)))

    def close(self) -> None:
        if self._h:
            lib().gt_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req: Request) -> int:
        t = C.c_longlong()
        _check(lib().gt_server_submit(self._h, C.byref(req), C.byref(t)))
        return t.value

    def wait(self, ticket: int, check: bool = True) -> dict:
        """The request's result; with check=False a failed request's status
        is returned in the dict instead of raised."""
        r = RequestResult()
        rc = lib().gt_server_wait(self._h, ticket, C.byref(r))
        if check:
            _check(rc)
        return {k: getattr(r, k) for k, _ in RequestResult._fields_}

    def cancel(self, ticket: int) -> None:
        """gt_server_cancel: GTError(GT_ESTATE) once a worker has taken it."""
        _check(lib().gt_server_cancel(self._h, ticket))

    def replay(self, reqs, arrival_ms) -> list:
        n = len(reqs)
        arr = (Request * n)(*reqs)
        t = (C.c_double * n)(*[float(x) for x in arrival_ms])
        res = (RequestResult * n)()
        rc = lib().gt_server_replay(self._h, n, arr, t, res)
        out = [{k: getattr(r, k) for k, _ in RequestResult._fields_} for r in res]
        if rc != 0:
            raise GTError(rc, lib().gt_last_error().decode())
        return out


def tmp_info(path: str):
    """((width, height, frames, channels), type code) of a .tmp file."""
    dims = (C.c_int * 4)()
    t = C.c_int()
    _check(lib().gt_tmp_info(os.fsencode(path), dims, C.byref(t)))
    return tuple(dims), t.value


def tmp_read(path: str) -> np.ndarray:
    """FileTMP::load on the host: a (C, F, H, W) float32 array."""
    W, H, F, Cc = tmp_info(path)[0]
    out = np.empty((Cc, F, H, W), np.float32)
    dims = (C.c_int * 4)()
    _check(lib().gt_tmp_read(os.fsencode(path), out.ctypes.data, out.size, dims))
    return out


def tmp_write(path: str, im: np.ndarray, type_name: str = "float32") -> None:
    """FileTMP::save on the host from a (C, F, H, W) float32 array."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    Cc, F, H, W = im.shape
    _check(lib().gt_tmp_write(os.fsencode(path), im.ctypes.data, W, H, F, Cc, TMP_TYPES[type_name]))


class NetServer:
    """gt_net_server: the -send / -receive protocol front end.  op is
    GT_NET_ECHO or GT_NET_BILATERAL; reply_port 0 answers on the same
    connection."""

    def __init__(self, op: int = GT_NET_BILATERAL, port: int = 0, reply_port: int = 0, sigma_s: float = 8.0,
                 sigma_r: float = 0.1, storage: str = "f32", max_pixels: int = 0, device: int = 0, slots: int = 4,
                 max_batch: int = 16, warm_side: int = 0, max_bytes: int = 0, max_inflight: int = 0,
                 bind_addr: str | None = None):
        self._h = _vp()
        cfg = NetConfig(port, reply_port, op, sigma_s, sigma_r, STORAGE[storage], max_pixels,
                        ServerConfig(slots, max_batch, 1 << 22, 0, warm_side), max_bytes, max_inflight,
                        bind_addr.encode() if bind_addr else None)
        _check(lib().gt_net_server_create(C.byref(self._h), device, C.byref(cfg)))
        self.port = lib().gt_net_server_port(self._h)

    def stats(self) -> dict:
        st = NetStats()
        _check(lib().gt_net_server_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in NetStats._fields_}

    def close(self) -> None:
        if self._h:
            lib().gt_net_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def send_image(sock, im: np.ndarray) -> None:
    """TCPConnection::sendImage: header (W, H, F, C) uint32, then planar float32."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    Cc, F, H, W = im.shape
    sock.sendall(np.array([W, H, F, Cc], np.uint32).tobytes() + im.tobytes())


def recv_image(sock) -> np.ndarray:
    """TCPConnection::recvImage: the inverse of send_image."""
    def take(k):
        buf = bytearray()
        while len(buf) < k:
            chunk = sock.recv(min(1 << 20, k - len(buf)))
            if not chunk:
                raise ConnectionError("connection closed mid-image")
            buf += chunk
        return bytes(buf)
    W, H, F, Cc = np.frombuffer(take(16), np.uint32)
    return np.frombuffer(take(4 * int(W) * int(H) * int(F) * int(Cc)), np.float32).reshape(int(Cc), int(F), int(H),
                                                                                         int(W)).copy()
