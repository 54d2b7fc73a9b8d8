"""The bilateral grid (GRID method) on the GPU against the oracle: the FP64
grid of the same cells within the 1e-5 gate, the golden reference outputs,
a distinct slice, the zero-sigma recursion, and the plugin's AUTO choice for
grey images (d = 3 <= 4 position channels picks GRID,
src/GaussTransform.cpp:553-564)."""
import os

import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("W,H,C,ss,sr", [(96, 64, 1, 4.0, 0.1), (80, 60, 3, 6.0, 0.2), (257, 129, 1, 8.0, 0.05)])
def test_joint_bilateral_grid_vs_fp64(gpu_ctx, oracle_lib, W, H, C, ss, sr):
    print(f"\n[case] grid bilateral {W}x{H}x{C} ss={ss} sr={sr}", flush=True)
    im = S.rect_image(W, H, C, K=8, noise=0.05, seed=W + C)
    out = gpu_ctx.joint_bilateral_grid(dev(im), dev(im), ss, ss, 0.0, sr).cpu().numpy()
    o64 = oracle_lib.joint_bilateral_grid(im, im, ss, ss, 0.0, sr, mode=oracle_lib.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL


def test_grid_golden(gpu_ctx):
    print("\n[case] grid golden cases", flush=True)
    f = np.load(os.path.join(GOLD, "grid_cases.npz"), allow_pickle=False)
    g = gpu_ctx.joint_bilateral_grid(dev(f["grey"]), dev(f["grey"]), 4.0, 4.0, 0.0, 0.1).cpu().numpy()
    assert rel_err(g, f["out_grey"]) < 2 * FP32_REL_TOL
    r = gpu_ctx.joint_bilateral_grid(dev(f["rgb"]), dev(f["rgb"]), 3.0, 3.0, 0.0, 0.2).cpu().numpy()
    assert rel_err(r, f["out_rgb"]) < 2 * FP32_REL_TOL
    inv = (np.float32(1) / f["sigmas"]).astype(np.float32)
    d = f["pos"].shape[0]
    out = gpu_ctx.gauss_transform_grid(dev(f["pos"].reshape(d, -1)), dev(f["values"].reshape(2, -1)), inv,
                                       slice_pos=dev(f["slice"].reshape(d, -1))).cpu().numpy()
    assert rel_err(out, f["out_gt"].reshape(2, -1)) < 2 * FP32_REL_TOL
    out = gpu_ctx.gauss_transform_grid(dev(f["pos"].reshape(d, -1)), dev(f["values"].reshape(2, -1)),
                                       inv).cpu().numpy()
    assert rel_err(out, f["out_gt_replay"].reshape(2, -1)) < 2 * FP32_REL_TOL


def test_grid_distinct_slice_vs_fp64(gpu_ctx, oracle_lib):
    print("\n[case] grid gauss transform, distinct slice, d=4", flush=True)
    rng = np.random.default_rng(9)
    pos = rng.normal(size=(4, 3000)).astype(np.float32)
    slc = (rng.normal(size=(4, 2000)) * 1.3).astype(np.float32)
    vals = np.concatenate([rng.uniform(size=(2, 3000)), np.ones((1, 3000))]).astype(np.float32)
    inv = np.array([1.5, 1.0, 0.7, 2.0], np.float32)
    out = gpu_ctx.gauss_transform_grid(dev(pos), dev(vals), inv, slice_pos=dev(slc)).cpu().numpy()
    o64, _ = oracle_lib.grid_transform(pos, vals, inv, slice_pos=slc, mode=oracle_lib.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL


def test_grid_zero_sigma_rows(gpu_ctx, oracle_lib):
    print("\n[case] grid bilateral sigma_y = 0", flush=True)
    im = S.rect_image(40, 12, 1, K=6, noise=0.05, seed=21)
    out = gpu_ctx.joint_bilateral_grid_host(im, im, 4.0, 0.0, 0.0, 0.2)
    o64 = oracle_lib.joint_bilateral_grid(im, im, 4.0, 0.0, 0.0, 0.2, mode=oracle_lib.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL


def test_grid_errors(gpu_ctx):
    from imagestack_amd import gtlattice
    print("\n[case] grid error paths", flush=True)
    im = S.rect_image(16, 16, 1, K=2, noise=0.05, seed=1)
    bad = im.copy()
    bad[0, 0, 3, 3] = np.nan
    with pytest.raises(gtlattice.GTError):
        gpu_ctx.joint_bilateral_grid_host(bad, bad, 4.0, 4.0, 0.0, 0.1)
    pos = np.zeros((7, 64), np.float32)
    with pytest.raises(gtlattice.GTError):
        gpu_ctx.gauss_transform_grid(dev(pos), dev(np.ones((1, 64), np.float32)), np.ones(7, np.float32))
    out = gpu_ctx.joint_bilateral_grid_host(im, im, 4.0, 4.0, 0.0, 0.1)   # context still usable
    assert np.isfinite(out).all()


def test_grid_nonfinite_values(gpu_ctx, oracle_lib):
    """NaN and +-Inf image values through the grid (JointBilateral GRID):
    NaN / +Inf / -Inf pixels exactly where the reference puts them (the F32
    grid oracle, bit-identical to the reference on such input), finite pixels
    within 1e-5 of the FP64 grid oracle; the next finite call is unaffected."""
    import torch
    O = oracle_lib
    print("\n[case] grid with NaN/Inf values", flush=True)
    guide = S.rect_image(90, 70, 1, K=8, noise=0.05, seed=61)
    im = S.uniform_image(90, 70, 2, seed=62)
    im[0, 0, 5, 7] = np.nan
    im[1, 0, 30, 40] = np.inf
    im[0, 0, 50, 60] = -np.inf
    im[1, 0, 10, 10] = np.inf
    im[1, 0, 10, 11] = -np.inf
    out = gpu_ctx.joint_bilateral_grid(dev(im), dev(guide), 3.0, 3.0, 0.0, 0.15).cpu().numpy()
    o32 = O.joint_bilateral_grid(im, guide, 3.0, 3.0, 0.0, 0.15)
    o64 = O.joint_bilateral_grid(im, guide, 3.0, 3.0, 0.0, 0.15, mode=O.F64_LAT)
    for f in (np.isnan, np.isposinf, np.isneginf):
        assert np.array_equal(f(out), f(o32)), f.__name__
    assert np.isnan(out).any() and np.isinf(out).any()
    fin = np.isfinite(o64)
    assert np.array_equal(fin, np.isfinite(out))
    assert rel_err(out[fin], o64[fin]) < FP32_REL_TOL
    clean = S.uniform_image(90, 70, 2, seed=62)
    a = gpu_ctx.joint_bilateral_grid(dev(clean), dev(guide), 3.0, 3.0, 0.0, 0.15).cpu().numpy()
    assert np.isfinite(a).all()
    assert rel_err(a, O.joint_bilateral_grid(clean, guide, 3.0, 3.0, 0.0, 0.15, mode=O.F64_LAT)) < FP32_REL_TOL
