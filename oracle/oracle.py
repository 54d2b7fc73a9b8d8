"""ctypes bindings for the test oracle (liboracle.so, our CPU restatement of
the reference lattice) and for the reference itself built from its sources
(oracle/_ref/libref_*.so).  TEST INFRASTRUCTURE ONLY: imported by tests/, the
smoke check and bench.py's cpu_baseline leg to check results, never by the
product path.

Arrays are dense planar numpy float32 images of shape (C, F, H, W) or (C, N).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F32, F64_LAT, F64_PURE = 0, 1, 2

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_sp = C.POINTER(C.c_short)
_ip = C.POINTER(C.c_int)
_lib = None
_ref = {}


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-C", HERE, "-j8", "all"], check=True,
                   stdout=subprocess.DEVNULL)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_scale_factors.argtypes = [C.c_int, _fp]
        L.orc_simplex.argtypes = [C.c_int, _fp, _sp, _fp, _ip, _ip]
        L.orc_simplex_batch.argtypes = [C.c_int, C.c_int, _fp, C.c_int64, _fp, _sp, _fp]
        L.orc_gauss_transform.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int64,
                                          _fp, C.c_int64, C.c_int, _fp, C.c_int64, _fp,
                                          C.c_void_p, C.c_int64, _ip]
        L.orc_lattice_build.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int64,
                                        _fp, C.c_int64, _fp]
        L.orc_lattice_build.restype = C.c_void_p
        L.orc_lattice_size.argtypes = [C.c_void_p]
        L.orc_lattice_export.argtypes = [C.c_void_p, _sp, _ip, _dp, _dp]
        L.orc_lattice_free.argtypes = [C.c_void_p]
        L.orc_joint_bilateral.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp,
                                          C.c_int, _fp, C.c_float, C.c_float, C.c_float,
                                          C.c_float, C.c_void_p]
        L.orc_joint_bilateral_grid.argtypes = L.orc_joint_bilateral.argtypes
        L.orc_grid_transform.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int64, _fp, C.c_int64,
                                         C.c_int, _fp, C.c_int64, _fp, C.c_void_p, C.c_int64]
        L.orc_grid_transform.restype = C.c_int64
        L.orc_patch_project.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int, C.c_int,
                                        C.c_int, C.c_int, _fp, C.c_void_p]
        L.orc_nlmeans.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int, C.c_int, C.c_int,
                                  C.c_int, _fp, C.c_float, C.c_float, C.c_void_p]
        _lib = L
    return _lib


def _planar(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(a.shape[0], -1)


def scale_factors(d: int) -> np.ndarray:
    sf = np.zeros(d, np.float32)
    lib().orc_scale_factors(d, _ptr(sf, _fp))
    return sf


def simplex(pos) -> tuple:
    """Enclosing simplex of one scaled position: (keys (d+1,d) int16,
    weights (d+1) f32, greedy, rank, overflow)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    d = pos.size
    keys = np.zeros((d + 1, d), np.int16)
    w = np.zeros(d + 1, np.float32)
    g = np.zeros(d + 1, np.int32)
    r = np.zeros(d + 1, np.int32)
    ovf = lib().orc_simplex(d, _ptr(pos, _fp), _ptr(keys, _sp), _ptr(w, _fp), _ptr(g, _ip), _ptr(r, _ip))
    return keys, w, g, r, ovf


def simplex_batch(splat, inv_sigma) -> tuple:
    """Per-pixel vertex keys (N, d+1, d) int16 and weights (N, d+1) f32."""
    sp = _planar(splat)
    d, n = sp.shape
    inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
    keys = np.zeros((n, d + 1, d), np.int16)
    w = np.zeros((n, d + 1), np.float32)
    rc = lib().orc_simplex_batch(n, d, _ptr(sp, _fp), n, _ptr(inv, _fp), _ptr(keys, _sp), _ptr(w, _fp))
    if rc < 0:
        raise ValueError("bad dimensions")
    return keys, w


def gauss_transform(splat, values, inv_sigma, slice_pos=None, mode: int = F32):
    """GaussTransform::apply(PERMUTOHEDRAL) restated.  Returns (out (vd, N), M)."""
    sp = _planar(splat)
    va = _planar(values)
    d, n = sp.shape
    vd = va.shape[0]
    inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
    if slice_pos is not None:
        sl = _planar(slice_pos)
        ns = sl.shape[1]
        slp = _ptr(sl, _fp)
    else:
        sl, ns, slp = None, n, None
    out = np.zeros((vd, ns), np.float32 if mode == F32 else np.float64)
    M = C.c_int(0)
    rc = lib().orc_gauss_transform(mode, n, d, vd, _ptr(sp, _fp), n, _ptr(va, _fp), n, ns, slp, ns,
                                   _ptr(inv, _fp), out.ctypes.data_as(C.c_void_p), ns, C.byref(M))
    if rc < 0:
        raise ValueError("bad dimensions")
    return out, M.value


def lattice(splat, values, inv_sigma, mode: int = F32) -> dict:
    """Sorted lattice: keys (M,d) int16, nbr (d+1, M, 2) int32, splat and blur
    values (M, vd) float64."""
    sp = _planar(splat)
    va = _planar(values)
    d, n = sp.shape
    vd = va.shape[0]
    inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
    L = lib()
    h = L.orc_lattice_build(mode, n, d, vd, _ptr(sp, _fp), n, _ptr(va, _fp), n, _ptr(inv, _fp))
    if not h:
        raise ValueError("bad dimensions")
    try:
        M = L.orc_lattice_size(h)
        keys = np.zeros((M, d), np.int16)
        nbr = np.zeros((d + 1, M, 2), np.int32)
        sv = np.zeros((M, vd), np.float64)
        bv = np.zeros((M, vd), np.float64)
        L.orc_lattice_export(h, _ptr(keys, _sp), _ptr(nbr, _ip), _ptr(sv, _dp), _ptr(bv, _dp))
    finally:
        L.orc_lattice_free(h)
    return {"M": M, "keys": keys, "nbr": nbr, "splat": sv, "blur": bv}


def joint_bilateral(im, ref, fw, fh, ff, color_sigma, mode: int = F32):
    """JointBilateral::apply(PERMUTOHEDRAL) restated on (C,F,H,W) images."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    Cc, F, H, W = im.shape
    Cr = ref.shape[0]
    out = np.zeros(im.shape, np.float32 if mode == F32 else np.float64)
    rc = lib().orc_joint_bilateral(mode, W, H, F, Cc, _ptr(im, _fp), Cr, _ptr(ref, _fp),
                                   fw, fh, ff, color_sigma, out.ctypes.data_as(C.c_void_p))
    if rc < 0:
        raise ValueError("bad arguments")
    return out


def grid_transform(splat, values, inv_sigma, slice_pos=None, mode: int = F32):
    """GaussTransform::apply(GRID) restated (src/DenseGrid.h, taps 5).
    Returns (out (vd, Ns), grid cells)."""
    sp = _planar(splat)
    va = _planar(values)
    d, n = sp.shape
    vd = va.shape[0]
    inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
    if slice_pos is not None:
        sl = _planar(slice_pos)
        ns = sl.shape[1]
        slp = _ptr(sl, _fp)
    else:
        sl, ns, slp = None, n, None
    out = np.zeros((vd, ns), np.float32 if mode == F32 else np.float64)
    cells = lib().orc_grid_transform(mode, n, d, vd, _ptr(sp, _fp), n, _ptr(va, _fp), n, ns, slp, ns,
                                     _ptr(inv, _fp), out.ctypes.data_as(C.c_void_p), ns)
    if cells < 0:
        raise ValueError("bad dimensions")
    return out, cells


def joint_bilateral_grid(im, ref, fw, fh, ff, color_sigma, mode: int = F32):
    """JointBilateral::apply(GRID) restated on (C,F,H,W) images."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    Cc, F, H, W = im.shape
    out = np.zeros(im.shape, np.float32 if mode == F32 else np.float64)
    rc = lib().orc_joint_bilateral_grid(mode, W, H, F, Cc, _ptr(im, _fp), ref.shape[0], _ptr(ref, _fp),
                                        fw, fh, ff, color_sigma, out.ctypes.data_as(C.c_void_p))
    if rc < 0:
        raise ValueError("bad arguments")
    return out


def patch_project(im, filters, mode: int = F32):
    """Convolve::apply(im, filters, Zero, Inner) restated: im (C,F,H,W),
    filters (K*C, ft, fh, fw) -> (K,F,H,W)."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    filt = np.ascontiguousarray(filters, dtype=np.float32)
    Cc, F, H, W = im.shape
    KC, ft, fh, fw = filt.shape
    if KC % Cc:
        raise ValueError("filter channels must be a multiple of the image channels")
    K = KC // Cc
    out = np.zeros((K, F, H, W), np.float32 if mode == F32 else np.float64)
    rc = lib().orc_patch_project(mode, W, H, F, Cc, _ptr(im, _fp), fw, fh, ft, K, _ptr(filt, _fp),
                                 out.ctypes.data_as(C.c_void_p))
    if rc < 0:
        raise ValueError("bad arguments")
    return out


def nlmeans(im, filters, spatial_sigma, patch_sigma, mode: int = F32):
    """NLMeans::apply after the PCA (FP32 guide, then the joint bilateral
    in `mode`).  Raises NotImplementedError where AUTO picks EXACT/GKDTREE."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    filt = np.ascontiguousarray(filters, dtype=np.float32)
    Cc, F, H, W = im.shape
    KC, ft, fh, fw = filt.shape
    out = np.zeros(im.shape, np.float32 if mode == F32 else np.float64)
    rc = lib().orc_nlmeans(mode, W, H, F, Cc, _ptr(im, _fp), fw, fh, ft, KC // Cc, _ptr(filt, _fp),
                           spatial_sigma, patch_sigma, out.ctypes.data_as(C.c_void_p))
    if rc == -2:
        raise NotImplementedError("JointBilateral's AUTO choice is EXACT or GKDTREE")
    if rc < 0:
        raise ValueError("bad arguments")
    return out


# --- the reference, compiled from its own sources (oracle/_ref) -------------

def ref_available(kind: str = "strict") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libref_{kind}.so"))


def ref_lib(kind: str = "strict"):
    if kind not in _ref:
        path = os.path.join(HERE, "_ref", f"libref_{kind}.so")
        L = C.CDLL(path)
        L.ref_gauss_transform.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp,
                                          _fp, _fp, _fp]
        L.ref_joint_bilateral.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp,
                                          C.c_float, C.c_float, C.c_float, C.c_float, _fp]
        L.ref_lattice.argtypes = [C.c_int, C.c_int, C.c_int, _fp, _fp, _fp, C.c_int, C.c_int,
                                  _sp, _fp]
        L.ref_grid_transform.argtypes = L.ref_gauss_transform.argtypes
        L.ref_joint_bilateral_grid.argtypes = L.ref_joint_bilateral.argtypes
        L.ref_convolve_inner.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_int, C.c_int, C.c_int,
                                         C.c_int, _fp, _fp]
        L.ref_patch_pca.argtypes = [C.c_uint, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_float,
                                    C.c_int, C.c_longlong, _fp]
        L.ref_nlmeans.argtypes = [C.c_uint, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_float, C.c_int,
                                  C.c_float, C.c_float, _fp]
        _ref[kind] = L
    return _ref[kind]


def ref_gauss_transform(splat, values, sigmas, slice_pos=None, kind: str = "strict"):
    """Reference GaussTransform::apply(PERMUTOHEDRAL) on (C,F,H,W) images."""
    sp = np.ascontiguousarray(splat, dtype=np.float32)
    va = np.ascontiguousarray(values, dtype=np.float32)
    d, F, H, W = sp.shape
    vd = va.shape[0]
    sig = np.ascontiguousarray(sigmas, dtype=np.float32)
    out = np.zeros((vd, F, H, W), np.float32)
    sl = None if slice_pos is None else np.ascontiguousarray(slice_pos, dtype=np.float32)
    rc = ref_lib(kind).ref_gauss_transform(W, H, F, d, _ptr(sp, _fp), vd, _ptr(va, _fp),
                                           None if sl is None else _ptr(sl, _fp), _ptr(sig, _fp),
                                           _ptr(out, _fp))
    if rc != 0:
        raise RuntimeError("reference raised")
    return out


def ref_joint_bilateral(im, ref, fw, fh, ff, color_sigma, kind: str = "strict", grid: bool = False):
    im = np.ascontiguousarray(im, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    Cc, F, H, W = im.shape
    out = np.zeros(im.shape, np.float32)
    L = ref_lib(kind)
    fn = L.ref_joint_bilateral_grid if grid else L.ref_joint_bilateral
    rc = fn(W, H, F, Cc, _ptr(im, _fp), ref.shape[0], _ptr(ref, _fp), fw, fh, ff, color_sigma, _ptr(out, _fp))
    if rc != 0:
        raise RuntimeError("reference raised")
    return out


def ref_grid_transform(splat, values, sigmas, slice_pos=None, kind: str = "strict"):
    """Reference GaussTransform::apply(GRID) on (C,F,H,W) images."""
    sp = np.ascontiguousarray(splat, dtype=np.float32)
    va = np.ascontiguousarray(values, dtype=np.float32)
    d, F, H, W = sp.shape
    vd = va.shape[0]
    sig = np.ascontiguousarray(sigmas, dtype=np.float32)
    out = np.zeros((vd, F, H, W), np.float32)
    sl = None if slice_pos is None else np.ascontiguousarray(slice_pos, dtype=np.float32)
    rc = ref_lib(kind).ref_grid_transform(W, H, F, d, _ptr(sp, _fp), vd, _ptr(va, _fp),
                                          None if sl is None else _ptr(sl, _fp), _ptr(sig, _fp),
                                          _ptr(out, _fp))
    if rc != 0:
        raise RuntimeError("reference raised")
    return out


def ref_lattice(splat, values, inv_sigma, blur: bool = False, kind: str = "strict"):
    """Reference lattice after splat (and blur): insertion-ordered keys (M,d)
    int16 and values (M,vd) f32."""
    sp = _planar(splat)
    va = _planar(values)
    d, n = sp.shape
    vd = va.shape[0]
    inv = np.ascontiguousarray(inv_sigma, dtype=np.float32)
    L = ref_lib(kind)
    M = L.ref_lattice(n, d, vd, _ptr(sp, _fp), _ptr(va, _fp), _ptr(inv, _fp), int(blur), 0, None, None)
    keys = np.zeros((M, d), np.int16)
    vals = np.zeros((M, vd), np.float32)
    L.ref_lattice(n, d, vd, _ptr(sp, _fp), _ptr(va, _fp), _ptr(inv, _fp), int(blur), M,
                  _ptr(keys, _sp), _ptr(vals, _fp))
    return keys, vals


def ref_convolve_inner(im, filters, kind: str = "strict"):
    """Reference Convolve::apply(im, filters, Zero, Inner): (C,F,H,W) x
    (Kc,ft,fh,fw) -> (Kc/C,F,H,W)."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    filt = np.ascontiguousarray(filters, dtype=np.float32)
    Cc, F, H, W = im.shape
    KC, ft, fh, fw = filt.shape
    out = np.zeros((max(KC // Cc, 1), F, H, W), np.float32)
    rc = ref_lib(kind).ref_convolve_inner(W, H, F, Cc, _ptr(im, _fp), fw, fh, ft, KC, _ptr(filt, _fp),
                                          _ptr(out, _fp))
    if rc < 0:
        raise RuntimeError("reference raised")
    return out


def ref_patch_pca(im, sigma, dims, seed: int, three_d: bool = False, kind: str = "strict"):
    """Reference PatchPCA::apply (or PatchPCA3D) after srand(seed): filters
    (C*dims, ft, side, side)."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    Cc, F, H, W = im.shape
    L = ref_lib(kind)
    side = L.ref_patch_pca(seed, int(three_d), W, H, F, Cc, _ptr(im, _fp), sigma, dims, 0, None)
    if side < 0:
        raise RuntimeError("reference raised")
    ft = side if three_d else 1
    filt = np.zeros((Cc * dims, ft, side, side), np.float32)
    L.ref_patch_pca(seed, int(three_d), W, H, F, Cc, _ptr(im, _fp), sigma, dims, filt.size, _ptr(filt, _fp))
    return filt


def ref_nlmeans(im, sigma, dims, spatial_sigma, patch_sigma, seed: int, three_d: bool = False,
                kind: str = "strict"):
    """Reference NLMeans::apply (or NLMeans3D) after srand(seed)."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    Cc, F, H, W = im.shape
    out = np.zeros(im.shape, np.float32)
    rc = ref_lib(kind).ref_nlmeans(seed, int(three_d), W, H, F, Cc, _ptr(im, _fp), sigma, dims, spatial_sigma,
                                   patch_sigma, _ptr(out, _fp))
    if rc != 0:
        raise RuntimeError("reference raised")
    return out


def sort_keys(keys: np.ndarray) -> np.ndarray:
    """Row indices that sort int16 key rows lexicographically (signed)."""
    if keys.shape[0] == 0:
        return np.zeros(0, np.int64)
    return np.lexsort(keys.T[::-1])
