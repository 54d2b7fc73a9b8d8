"""The host key-packing plan (imagestack_amd/csrc/gt_plan.h), compiled with g++
on the CPU: every vertex key of the oracle's simplex lies inside the plan's
ranges, packed words sort in the lexicographic order of the keys, and the
plan's scale factors are bit-identical to the oracle's."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from imagestack_amd import synth as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = r'''
#include "gt_plan.h"
extern "C" void plan(int d, const float *lo, const float *hi, int *klo, unsigned *span, int *shift,
                     int *packed, int *bits, float *sf_out) {
    float sf[16];
    gt::plan_scale_factors(d, sf);
    gt::KeyPlan kp;
    gt::plan_from_bounds(d, sf, lo, hi, kp);
    for (int i = 0; i < d; i++) { klo[i] = kp.lo[i]; span[i] = kp.span[i]; shift[i] = kp.shift[i]; sf_out[i] = sf[i]; }
    *packed = kp.packed;
    *bits = kp.bits;
}
'''


@pytest.fixture(scope="module")
def plan_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("plan")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "libplan.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "imagestack_amd", "csrc"), str(src), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    vp = C.c_void_p
    lib.plan.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), vp]
    return lib


def make_plan(lib, pos):
    d = pos.shape[0]
    lo = np.ascontiguousarray(pos.reshape(d, -1).min(axis=1), np.float32)
    hi = np.ascontiguousarray(pos.reshape(d, -1).max(axis=1), np.float32)
    klo = np.zeros(d, np.int32)
    span = np.zeros(d, np.uint32)
    shift = np.zeros(d, np.int32)
    sf = np.zeros(d, np.float32)
    packed, bits = C.c_int(), C.c_int()
    lib.plan(d, lo.ctypes.data, hi.ctypes.data, klo.ctypes.data, span.ctypes.data, shift.ctypes.data,
             C.byref(packed), C.byref(bits), sf.ctypes.data)
    return klo, span, shift, packed.value, bits.value, sf


CASES = {
    "bilateral d=5": lambda: S.bilateral_positions(S.rect_image(96, 64, 3, K=8, noise=0.05, seed=3), 8.0, 0.1),
    "bilateral d=3 gray": lambda: S.bilateral_positions(S.rect_image(80, 60, 1, K=8, noise=0.05, seed=4), 4.0, 0.05),
    "highdim d=16": lambda: S.highdim_positions(24, 24, nfeat=14, sigma_s=12.0, seed=5),
    "wide ranges d=4": lambda: (np.random.default_rng(0).normal(size=(4, 1, 40, 40)) * 300).astype(np.float32),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_keys_inside_plan_and_packed_order(plan_lib, oracle_lib, name):
    pos = CASES[name]()
    d = pos.shape[0]
    klo, span, shift, packed, bits, sf = make_plan(plan_lib, pos)
    assert np.array_equal(sf, oracle_lib.scale_factors(d))
    keys, _ = oracle_lib.simplex_batch(pos.reshape(d, -1), np.ones(d, np.float32))
    k = keys.reshape(-1, d).astype(np.int64)
    rel = k - klo[None, :]
    assert (rel >= 0).all() and (rel <= span[None, :].astype(np.int64)).all()
    assert bits == sum(int(s).bit_length() for s in span)
    if packed:
        assert bits <= 63
        words = (rel.astype(np.uint64) << shift[None, :].astype(np.uint64)).sum(axis=1, dtype=np.uint64)
        uniq = np.unique(k, axis=0)
        uw = ((uniq - klo).astype(np.uint64) << shift.astype(np.uint64)).sum(axis=1, dtype=np.uint64)
        # unsigned order of the words == lexicographic order of the keys
        assert np.array_equal(np.argsort(uw, kind="stable"), oracle_lib.sort_keys(uniq.astype(np.int16)))
        assert len(np.unique(words)) == uniq.shape[0]


def test_plan_is_tight_for_bilateral(plan_lib):
    """The cfg-like 5-D bilateral plan packs into one word."""
    pos = CASES["bilateral d=5"]()
    _, _, _, packed, bits, _ = make_plan(plan_lib, pos)
    assert packed == 1 and bits <= 63
