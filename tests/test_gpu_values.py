"""Value handling of the lattice on the GPU.

  * The fixed-point scales of the splat sums come from each call's own value
    bounds, so a call gives the same bits whatever ran on the context before
    it (a wide-range image followed by a [0, 1] one, in every form).
  * High-dynamic-range values (seven decades) stay within 1e-5 of the FP64
    oracle per element, with a floor of 1e-9 of the largest magnitude.
  * NaN and +-Inf values spread as in the reference's FP32 arithmetic
    (src/Permutohedral.h:377 splat, :467-469 blur, :423 slice): the NaN, +Inf
    and -Inf pixels of the output are exactly the reference's (the F32
    oracle's, bit-identical to the reference on such input), and every finite
    pixel is within 1e-5 of the FP64 oracle.  Covered for the LDS-tiled splat
    (d <= 10, vd <= 4), the direct splat (vd > 4 and d > 10), the batched
    lattice and both vertex-numbering forms.
"""
import numpy as np
import pytest

from imagestack_amd import gtlattice as G
from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def planar(a):
    a = np.ascontiguousarray(a, np.float32)
    return a.reshape(a.shape[0], -1)


def case(name):
    print(f"\n[case] {name}", flush=True)


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def check_nonfinite(out, o32, o64):
    out = np.asarray(out, np.float32).reshape(o32.shape)
    for f in (np.isnan, np.isposinf, np.isneginf):
        assert np.array_equal(f(out), f(o32)), f.__name__
    assert np.isnan(o32).any() or np.isinf(o32).any()
    fin = np.isfinite(o64)
    assert np.array_equal(fin, np.isfinite(out))
    assert rel_err(out[fin], o64[fin]) < FP32_REL_TOL


# --- history independence ----------------------------------------------------

def jb(ctx, im, guide, ss, sr):
    return ctx.joint_bilateral(dev(im), dev(guide), ss, ss, 0.0, sr).cpu().numpy()


def gt(ctx, pos, val, W):
    return ctx.gauss_transform(dev(pos), dev(val), np.ones(pos.shape[0], np.float32), row_w=W).cpu().numpy()


def test_scales_independent_of_earlier_calls(gpu_ctx):
    """A 1e4-range call, then a [0, 1] call of the same shape and filter: the
    second must equal a fresh context's output bit for bit (its scale is not
    the earlier call's coarser one)."""
    case("wide-range call, then a [0,1] call, on one context")
    guide = S.rect_image(320, 240, 3, K=24, noise=0.05, seed=61)
    small = S.uniform_image(320, 240, 3, seed=62)
    wide = (small * np.float32(1e4)).astype(np.float32)
    pos = planar(S.bilateral_positions(guide, 8.0, 0.1))
    vals_small = planar(small)
    vals_wide = planar(wide)
    ctx = G.Context(0)
    try:
        seq = [jb(ctx, wide, guide, 8.0, 0.1), jb(ctx, small, guide, 8.0, 0.1),
               gt(ctx, pos, vals_wide, 320), gt(ctx, pos, vals_small, 320)]
    finally:
        ctx.close()
    fresh = []
    for k, f in enumerate([lambda c: jb(c, wide, guide, 8.0, 0.1), lambda c: jb(c, small, guide, 8.0, 0.1),
                           lambda c: gt(c, pos, vals_wide, 320), lambda c: gt(c, pos, vals_small, 320)]):
        c = G.Context(0)
        try:
            fresh.append(f(c))
        finally:
            c.close()
    for k, (a, b) in enumerate(zip(seq, fresh)):
        assert same_bits(a, b), k


def test_batch_scales_independent_of_earlier_calls():
    import torch
    case("batch: wide-range images, then [0,1] images, on one context")
    imgs = S.batch_images(seed=9, count=6)
    small, wide = [], []
    for im, ss, sr in imgs:
        a = torch.from_numpy(im).cuda()[:, None]
        small.append((a, a, torch.empty_like(a), ss, sr))
        w = torch.from_numpy((im * np.float32(1e4)).astype(np.float32)).cuda()[:, None]
        wide.append((w, a, torch.empty_like(w), ss, sr))
    ctx = G.Context(0)
    try:
        ctx.joint_bilateral_batch(wide)
        ctx.joint_bilateral_batch(small)
        got = [o.cpu().numpy().copy() for (_, _, o, _, _) in small]
    finally:
        ctx.close()
    fresh = G.Context(0)
    try:
        fresh.joint_bilateral_batch(small)
        want = [o.cpu().numpy() for (_, _, o, _, _) in small]
    finally:
        fresh.close()
    for a, b in zip(got, want):
        assert same_bits(a, b)


# --- high dynamic range ------------------------------------------------------

def test_hdr_values_per_element(gpu_ctx, oracle_lib):
    O = oracle_lib
    case("HDR values, seven decades, joint bilateral and gauss transform 256x192")
    guide = S.rect_image(256, 192, 3, K=20, noise=0.05, seed=71)
    u = S.uniform_image(256, 192, 3, seed=72)
    hdr = np.power(np.float32(10.0), u * np.float32(7.0) - np.float32(3.0)).astype(np.float32)
    out = jb(gpu_ctx, hdr, guide, 6.0, 0.1)
    o64 = O.joint_bilateral(hdr, guide, 6.0, 6.0, 0.0, 0.1, mode=O.F64_LAT)
    e = rel_err(out, o64, floor=1e-9)
    print("joint bilateral per-element err", e, flush=True)
    assert e < FP32_REL_TOL
    pos = planar(S.bilateral_positions(guide, 6.0, 0.1))
    val = planar(S.homogeneous_values(hdr))
    out = gt(gpu_ctx, pos, val, 256)
    o64, _ = O.gauss_transform(pos, val, np.ones(5, np.float32), mode=O.F64_LAT)
    e = rel_err(out, o64.reshape(out.shape), floor=1e-9)
    print("gauss transform per-element err", e, flush=True)
    assert e < FP32_REL_TOL


# --- NaN and Inf ---------------------------------------------------------------

def plant(a, seed):
    """NaN, +Inf, -Inf and a +Inf/-Inf pair one pixel apart in channel planes."""
    a = np.array(a, np.float32, copy=True)
    flat = a.reshape(a.shape[0], -1)
    n = flat.shape[1]
    rng = np.random.default_rng(seed)
    idx = rng.choice(n - 1, 6, replace=False)
    flat[0, idx[0]] = np.nan
    flat[-1, idx[1]] = np.inf
    flat[0, idx[2]] = -np.inf
    flat[min(1, flat.shape[0] - 1), idx[3]] = np.inf
    flat[min(1, flat.shape[0] - 1), idx[3] + 1] = -np.inf
    return a


@pytest.mark.parametrize("neighbours", [1, 2])
def test_nonfinite_joint_bilateral(gpu_ctx, oracle_lib, neighbours):
    O = oracle_lib
    case(f"NaN/Inf image, joint bilateral 200x150 (tiled splat, neighbours form {neighbours})")
    guide = S.rect_image(200, 150, 3, K=12, noise=0.05, seed=81)
    im = plant(S.uniform_image(200, 150, 3, seed=82), 83)
    out = gpu_ctx.joint_bilateral(dev(im), dev(guide), 5.0, 5.0, 0.0, 0.15,
                                  tune={"neighbours": neighbours}).cpu().numpy()
    o32 = O.joint_bilateral(im, guide, 5.0, 5.0, 0.0, 0.15)
    o64 = O.joint_bilateral(im, guide, 5.0, 5.0, 0.0, 0.15, mode=O.F64_LAT)
    check_nonfinite(out, o32, o64)
    # the next call on the context, finite again, is as a fresh context's
    clean = S.uniform_image(200, 150, 3, seed=82)
    a = jb(gpu_ctx, clean, guide, 5.0, 0.15)
    c = G.Context(0)
    try:
        b = jb(c, clean, guide, 5.0, 0.15)
    finally:
        c.close()
    assert same_bits(a, b)


@pytest.mark.parametrize("d,vd", [(3, 2), (8, 4), (5, 6), (12, 3)])
def test_nonfinite_gauss_transform(gpu_ctx, oracle_lib, d, vd):
    """d = 3 and 8: the LDS-tiled splat; vd = 6 and d = 12: the direct one."""
    O = oracle_lib
    case(f"NaN/Inf values, gauss transform 96x80 d={d} vd={vd}")
    W, H = 96, 80
    s = S.Stream(90 + d)
    pos = (s.normal(d * W * H).reshape(d, -1) * 1.5).astype(np.float32)
    val = plant(s.uniform(vd * W * H).reshape(vd, -1).astype(np.float32), d)
    inv = np.ones(d, np.float32)
    out = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W).cpu().numpy()
    o32, _ = O.gauss_transform(pos, val, inv)
    o64, _ = O.gauss_transform(pos, val, inv, mode=O.F64_LAT)
    check_nonfinite(out, o32.reshape(out.shape), o64.reshape(out.shape))


def test_nonfinite_batch(oracle_lib):
    import torch
    O = oracle_lib
    case("NaN/Inf images in a batch")
    imgs = S.batch_images(seed=10, count=5)
    items, host = [], []
    for k, (im, ss, sr) in enumerate(imgs):
        bad = plant(im, 100 + k) if k % 2 == 0 else im
        g = torch.from_numpy(im).cuda()[:, None]
        a = torch.from_numpy(bad).cuda()[:, None]
        items.append((a, g, torch.empty_like(a), ss, sr))
        host.append((bad[:, None], im[:, None], ss, sr))
    ctx = G.Context(0)
    try:
        ctx.joint_bilateral_batch(items)
        outs = [o.cpu().numpy() for (_, _, o, _, _) in items]
    finally:
        ctx.close()
    for k, (out, (bad, im, ss, sr)) in enumerate(zip(outs, host)):
        o32 = O.joint_bilateral(bad, im, ss, ss, 0.0, sr)
        o64 = O.joint_bilateral(bad, im, ss, ss, 0.0, sr, mode=O.F64_LAT)
        if k % 2 == 0:
            check_nonfinite(out, o32, o64)
        else:
            assert np.all(np.isfinite(out)) and rel_err(out, o64) < FP32_REL_TOL


def test_more_than_16_value_channels(gpu_ctx, oracle_lib):
    """vd = 20 (the reference takes any number of value channels): run as
    groups of 16 over the same positions, the output equals the groups run
    one by one bit for bit, and the FP64 oracle of all 20 channels."""
    O = oracle_lib
    case("gauss transform d=5 vd=20")
    W, H = 80, 64
    guide = S.rect_image(W, H, 3, K=10, noise=0.05, seed=91)
    pos = planar(S.bilateral_positions(guide, 6.0, 0.1))
    val = S.Stream(92).uniform(20 * W * H).reshape(20, -1).astype(np.float32)
    inv = np.ones(5, np.float32)
    out = gt(gpu_ctx, pos, val, W)
    a = gt(gpu_ctx, pos, val[:16], W)
    b = gt(gpu_ctx, pos, val[16:], W)
    assert same_bits(out, np.concatenate([a, b]))
    o64, _ = O.gauss_transform(pos, val, inv, mode=O.F64_LAT)
    assert rel_err(out, o64.reshape(out.shape)) < FP32_REL_TOL
    case("joint bilateral with 17 image channels")
    im = S.uniform_image(W, H, 17, seed=93)
    got = jb(gpu_ctx, im, guide, 6.0, 0.1)
    o64 = O.joint_bilateral(im, guide, 6.0, 6.0, 0.0, 0.1, mode=O.F64_LAT)
    assert rel_err(got, o64) < FP32_REL_TOL
