"""Edge cases of the path on the GPU against the oracle: no Gaussians, a
single pixel, one-row and one-column images, ragged (prime) sizes, every
pixel at the same position (one simplex, all collisions), spatial sigmas too
large to filter (JointBilateral then drops the axis), and keys beyond int16
(GT_ERANGE, as the reference's short keys cannot hold them).  Non-finite
values are covered in test_gpu_values.py: they spread as in the reference."""
import numpy as np
import pytest

from imagestack_amd import gtlattice as G
from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_no_gaussians(gpu_ctx):
    print("\n[case] n = 0 splat points, 50 slice points", flush=True)
    import torch
    splat = torch.empty((3, 0), device="cuda")
    vals = torch.empty((2, 0), device="cuda")
    slc = dev(np.random.default_rng(1).normal(size=(3, 50)))
    out = gpu_ctx.gauss_transform(splat, vals, np.ones(3, np.float32), slice_pos=slc).cpu().numpy()
    assert out.shape == (2, 50) and np.all(out == 0)


@pytest.mark.parametrize("W,H,F,C,Cr,fw,fh,ff,cs", [
    (1, 1, 1, 3, 3, 4.0, 4.0, 0.0, 0.1),      # one pixel: only the colour axes
    (300, 1, 1, 3, 3, 6.0, 6.0, 0.0, 0.1),    # one row: y dropped
    (1, 200, 1, 1, 1, 5.0, 5.0, 0.0, 0.2),    # one column: x dropped
    (263, 131, 1, 3, 3, 7.0, 5.0, 0.0, 0.15), # ragged prime sizes, anisotropic sigmas
    (40, 30, 5, 2, 2, 3.0, 3.0, 2.0, 0.2),    # frames filtered (d = 5 with t)
    (64, 48, 1, 3, 3, 1e4, 6.0, 0.0, 0.1),    # sigma_x >= 10 W: x not a position channel
])
def test_joint_bilateral_shapes(gpu_ctx, oracle_lib, W, H, F, C, Cr, fw, fh, ff, cs):
    print(f"\n[case] joint bilateral {W}x{H}x{F} C={C} Cr={Cr} sigmas {fw},{fh},{ff} {cs}", flush=True)
    im = S.rect_image(W, H, C, K=6, noise=0.05, seed=W + H, F=F)
    ref = S.rect_image(W, H, Cr, K=6, noise=0.05, seed=W * H, F=F)
    out = gpu_ctx.joint_bilateral(dev(im), dev(ref), fw, fh, ff, cs).cpu().numpy()
    o64 = oracle_lib.joint_bilateral(im, ref, fw, fh, ff, cs, mode=oracle_lib.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL


def test_all_pixels_collide(gpu_ctx, oracle_lib):
    print("\n[case] constant guide: one simplex for every pixel", flush=True)
    W, H = 128, 96
    ref = np.full((3, 1, H, W), 0.37, np.float32)
    im = S.uniform_image(W, H, 3, seed=3)
    out = gpu_ctx.joint_bilateral(dev(im), dev(ref), 1e5, 1e5, 0.0, 0.1).cpu().numpy()
    o64 = oracle_lib.joint_bilateral(im, ref, 1e5, 1e5, 0.0, 0.1, mode=oracle_lib.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL
    # every pixel gets the same value: the image mean, up to the lattice's own weights
    assert np.ptp(out.reshape(3, -1), axis=1).max() < 1e-5
    pos = np.full((3, W * H), 0.37, np.float32)
    g = gpu_ctx.build_lattice(dev(pos), dev(np.ones((1, W * H), np.float32)), np.full(3, 10.0, np.float32))
    assert g["M"] == 4   # d + 1 vertices


@pytest.mark.parametrize("d", [3, 5])
def test_exact_ties_key_set(gpu_ctx, oracle_lib, d):
    """Positions on the half-integer grid of the lattice's own scale
    (a_i / (2 sf_i), a_i in -4..4): many elevate to exact ties -- greedy
    roundings halfway between multiples of d+1 (ties go down,
    src/Permutohedral.h:312) and equal differentials (the later index ranks
    higher, :321-324) -- which show as vertices of weight exactly 0.  The
    rank tie rule decides which zero-weight vertex is inserted: ranking ties
    to the earlier index gives another key set for these positions (checked
    with a restatement on the CPU), so the key set and the neighbour table
    must match the oracle bit for bit.  (The greedy tie direction is undone
    by the walk back onto the hyperplane here: both directions give the same
    simplex.)"""
    print(f"\n[case] exact-tie positions, d={d}", flush=True)
    O = oracle_lib
    sf = O.scale_factors(d).astype(np.float32)
    rng = np.random.default_rng(40 + d)
    n = 4096
    a = rng.integers(-4, 5, size=(d, n)).astype(np.float32)
    pos = (a / (np.float32(2.0) * sf[:, None])).astype(np.float32)
    inv = np.ones(d, np.float32)
    _, w = O.simplex_batch(pos, inv)
    ties = int((w == 0).any(axis=1).sum())
    print("pixels with a zero-weight vertex", ties, "of", n, flush=True)
    assert ties > n // 8
    val = np.ones((1, n), np.float32)
    g = gpu_ctx.build_lattice(dev(pos), dev(val), inv)
    o = O.lattice(pos, val, inv)
    assert np.array_equal(g["keys"], o["keys"])
    assert np.array_equal(g["nbr"], o["nbr"])


def test_keys_beyond_int16(gpu_ctx):
    print("\n[case] keys beyond int16", flush=True)
    pos = np.zeros((3, 64), np.float32)
    pos[1, 7] = 1e6
    val = np.ones((2, 64), np.float32)
    with pytest.raises(G.GTError) as e:
        gpu_ctx.gauss_transform(dev(pos), dev(val), np.ones(3, np.float32))
    assert e.value.code == 3
    pos[1, 7] = 0.0
    assert np.all(np.isfinite(gpu_ctx.gauss_transform(dev(pos), dev(val), np.ones(3, np.float32)).cpu().numpy()))
