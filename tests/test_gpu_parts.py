"""Row-band table parts (gt_tuning.parts): a one-frame bilateral form splits
the hash table into eight ranges of the last lattice coordinate (the row),
one per XCD.  Only where vertices sit in the table and which ids they get
change, so the output must equal the unsplit table's bit for bit, for both
neighbour forms, packed and wide keys, and match the FP64 oracle."""
import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


VARIANTS = [{}, {"parts": -1}, {"neighbours": 1}, {"neighbours": 2}, {"neighbours": 2, "parts": -1},
            {"neighbours": 1, "parts": -1}]


@pytest.mark.parametrize("W,H,Cr,ss,sr", [
    (320, 240, 3, 4.0, 0.1),      # d = 5, one-word keys
    (256, 200, 6, 6.0, 0.15),     # d = 8, wide keys
    (96, 700, 3, 2.0, 0.05),      # tall image: bands of few rows
    (700, 40, 1, 3.0, 0.1),       # wide, short image: d = 3
])
def test_parts_same_output(gpu_ctx, oracle_lib, W, H, Cr, ss, sr):
    print(f"\n[case] table parts {W}x{H} guide channels {Cr}", flush=True)
    ref = S.rect_image(W, H, Cr, K=24, noise=0.04, seed=W + Cr)
    im = S.uniform_image(W, H, 3, seed=H)
    outs = [gpu_ctx.joint_bilateral(dev(im), dev(ref), ss, ss, 0.0, sr, tune=t).cpu().numpy() for t in VARIANTS]
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    o64 = oracle_lib.joint_bilateral(im, ref, ss, ss, 0.0, sr, mode=oracle_lib.F64_LAT)
    assert rel_err(outs[0], o64) < FP32_REL_TOL


def test_parts_skip_volumes(gpu_ctx, oracle_lib):
    """Several frames filtered along t: the last position channel is the
    frame, not the row, so the table is not split; output as the oracle's."""
    print("\n[case] table parts skipped for a 3-frame volume", flush=True)
    im = S.rect_image(96, 80, 3, K=16, noise=0.04, seed=3, F=3)
    outs = [gpu_ctx.joint_bilateral(dev(im), dev(im), 4.0, 4.0, 1.0, 0.1, tune=t).cpu().numpy()
            for t in ({}, {"parts": -1})]
    assert np.array_equal(outs[0], outs[1])
    o64 = oracle_lib.joint_bilateral(im, im, 4.0, 4.0, 1.0, 0.1, mode=oracle_lib.F64_LAT)
    assert rel_err(outs[0], o64) < FP32_REL_TOL


def image_positions(W, H, nfeat, spread, seed):
    """nfeat smooth-ish feature channels scaled by `spread`, then x/8 and y/8
    (the row last): the layout whose table a Gauss transform splits."""
    im = S.rect_image(W, H, nfeat, K=40, noise=0.05, seed=seed)
    x = np.tile(np.arange(W, dtype=np.float32) / 8, (H, 1))
    y = np.tile((np.arange(H, dtype=np.float32) / 8)[:, None], (1, W))
    return np.concatenate([im.reshape(nfeat, -1) * spread, x.reshape(1, -1), y.reshape(1, -1)]).astype(np.float32)


@pytest.mark.parametrize("nfeat,spread,oracle", [(3, 4.0, True), (10, 2.0, False)])
def test_parts_gauss_transform(gpu_ctx, oracle_lib, nfeat, spread, oracle):
    """A Gauss transform over a 1024x1024 image (row_w given), rows in its
    last position channel: split table (d = 5 one-word keys, d = 12 wide
    keys), plain table and sorted ids give the same output bit for bit."""
    d = nfeat + 2
    print(f"\n[case] table parts, Gauss transform 1024x1024 d={d}", flush=True)
    W = H = 1024
    pos = image_positions(W, H, nfeat, spread, seed=d)
    rng = np.random.default_rng(d)
    val = np.concatenate([rng.uniform(size=(3, W * H)), np.ones((1, W * H))]).astype(np.float32)
    inv = np.ones(d, np.float32)
    outs = [gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune=t).cpu().numpy()
            for t in ({}, {"parts": -1}, {"neighbours": 1, "parts": -1})]
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    if oracle:
        o64, _ = oracle_lib.gauss_transform(pos, val, inv, mode=oracle_lib.F64_LAT)
        assert rel_err(outs[0], o64) < FP32_REL_TOL


def test_parts_skewed_last_channel(gpu_ctx):
    """A last position channel that is one value but for a single far pixel:
    every key of a dense lattice lands in the first part, the split table
    fills up, and the call is redone on the plain table; same output as
    without parts.  The failed attempt's neighbour and blur kernels walk the
    parts' id ranges, which must come out empty (M = 0) after a full table."""
    print("\n[case] table parts, skewed last channel (redo without parts)", flush=True)
    W = H = 1024
    pos = image_positions(W, H, 3, 40.0, seed=11)
    pos[-1] = 0.0
    pos[-1, 0] = 4000.0
    rng = np.random.default_rng(3)
    val = np.concatenate([rng.uniform(size=(2, W * H)), np.ones((1, W * H))]).astype(np.float32)
    inv = np.ones(5, np.float32)
    a, st = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, stats=True)
    b = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune={"parts": -1})
    print("retries", st["retries"], "M", st["M"])
    assert np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))
    assert st["retries"] >= 1


def test_parts_local_overflow_at_capacity_limit(gpu_ctx):
    """A split table whose capacity is already the most the call can need
    (M / n near d + 1, so n (d + 1) / load_factor slots), with a skewed last
    channel: one part overflows while the table as a whole has room.  The
    call is redone at the same capacity without parts rather than failing.
    The second call on the context starts at that capacity (the context
    remembers the vertex density).  Both equal the plain table bit for bit."""
    import torch
    from imagestack_amd import gtlattice as G
    print("\n[case] table parts overflow locally at the capacity limit", flush=True)
    W = H = 1024
    n = W * H
    rng = np.random.default_rng(17)
    pos = np.empty((3, n), np.float32)
    pos[:2] = rng.uniform(-2000.0, 2000.0, size=(2, n))   # nearly every vertex distinct: M ~ 4 n
    pos[2] = rng.normal(0.0, 0.3, size=n)                 # the last channel: one narrow band ...
    pos[2, 0], pos[2, 1] = -2000.0, 2000.0                # ... but bounds far apart
    val = np.concatenate([rng.uniform(size=(1, n)), np.ones((1, n))]).astype(np.float32)
    inv = np.ones(3, np.float32)
    ctx = G.Context(0)
    try:
        got = []
        for k in range(2):
            out, st = ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, stats=True)
            print("call", k, "M", st["M"], "capacity", st["capacity"], "retries", st["retries"], flush=True)
            got.append((out.cpu().numpy(), st))
        want = ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune={"parts": -1}).cpu().numpy()
    finally:
        ctx.close()
    assert got[0][1]["M"] > 2 * n
    assert got[1][1]["retries"] >= 1
    for out, _ in got:
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))


def test_table_fills_during_tiled_splat():
    """A first call whose table fills up while the tiled splat runs (2048^2
    RGB noise at sigma_s 4, sigma_r 0.1, about 5 vertices per pixel against a
    first table sized for 1): workgroups that start after the table is found
    full skip their flush as a whole, the call is redone in a larger table,
    and its output equals a call that starts at the right size."""
    from imagestack_amd import gtlattice as G
    print("\n[case] 2048^2 RGB noise, table full during the first splat", flush=True)
    im = S.uniform_image(2048, 2048, 3, seed=5)
    di = dev(im)
    ctx = G.Context(0)
    try:
        first, st0 = ctx.joint_bilateral(di, di, 4.0, 4.0, 0.0, 0.1, stats=True)
        first = first.cpu().numpy()
        again, st1 = ctx.joint_bilateral(di, di, 4.0, 4.0, 0.0, 0.1, stats=True)
        again = again.cpu().numpy()
    finally:
        ctx.close()
    print("first call M", st0["M"], "retries", st0["retries"], "then retries", st1["retries"], flush=True)
    assert st0["retries"] >= 1 and st1["retries"] == 0
    assert st0["M"] == st1["M"] > 2048 * 2048
    assert np.array_equal(first.view(np.uint32), again.view(np.uint32))
