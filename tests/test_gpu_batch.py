"""Batched lattices (gt_joint_bilateral_batch) and graph replays.

Each image of a batch is filtered as it would be alone; the splat
accumulates in fixed point, so a replayed graph reproduces its output bit for
bit and a batch matches the single-image runs to fixed-point rounding."""
import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def make_batch(seed=0, count=7, C=3):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        W = int(rng.integers(33, 200))
        H = int(rng.integers(17, 160))
        ref = S.rect_image(W, H, C, K=6, noise=0.05, seed=1000 + seed * 100 + k)
        im = S.uniform_image(W, H, C, seed=2000 + seed * 100 + k) if k % 2 else ref.copy()
        ss = float(rng.choice([4.0, 8.0, 16.0]))
        sr = float(rng.choice([0.05, 0.1, 0.2]))
        out.append((im, ref, ss, sr))
    return out


def test_batch_matches_single_and_oracle(gpu_ctx, oracle_lib):
    import torch
    print("\n[case] batch of 7 RGB images vs single runs and FP64 oracle", flush=True)
    batch = make_batch(1)
    items, dev_in = [], []
    for im, ref, ss, sr in batch:
        a, b = dev(im), dev(ref)
        o = torch.empty_like(a)
        items.append((a, b, o, ss, sr))
    st = gpu_ctx.joint_bilateral_batch(items, stats=True)
    print("batch M", st["M"], "sort passes", st["sort_passes"])
    for (im, ref, ss, sr), (a, b, o, _, _) in zip(batch, items):
        single = gpu_ctx.joint_bilateral(a, b, ss, ss, 0.0, sr).cpu().numpy()
        got = o.cpu().numpy()
        assert rel_err(got, single) < 1e-6
        o64 = oracle_lib.joint_bilateral(im, ref, ss, ss, 0.0, sr, mode=oracle_lib.F64_LAT)
        assert rel_err(got, o64) < FP32_REL_TOL


def test_batch_grayscale_narrow_tiles(gpu_ctx, oracle_lib):
    import torch
    print("\n[case] grayscale batch (d=3, vd=2), 16-pixel-wide tiles", flush=True)
    batch = make_batch(2, count=5, C=1)
    items = [(dev(im), dev(ref), torch.empty((1, 1) + im.shape[2:], device="cuda"), ss, sr)
             for im, ref, ss, sr in batch]
    gpu_ctx.joint_bilateral_batch(items, tune={"tile_w": 16})
    for (im, ref, ss, sr), (_, _, o, _, _) in zip(batch, items):
        o64 = oracle_lib.joint_bilateral(im, ref, ss, ss, 0.0, sr, mode=oracle_lib.F64_LAT)
        assert rel_err(o.cpu().numpy(), o64) < FP32_REL_TOL


def test_graph_replays_are_identical(gpu_ctx):
    """The captured pipeline graph replays three times, with eager single-image
    work and synchronisation between replays; every replay is bit-identical."""
    import torch
    print("\n[case] graph replays", flush=True)
    batch = make_batch(3, count=4)
    items = [(dev(im), dev(ref), None, ss, sr) for im, ref, ss, sr in batch]
    items = [(a, b, torch.empty_like(a), ss, sr) for a, b, _, ss, sr in items]
    outs = []
    for rep in range(3):
        for it in items:
            it[2].zero_()
        gpu_ctx.joint_bilateral_batch(items)
        outs.append([it[2].cpu().numpy() for it in items])
        eager = gpu_ctx.joint_bilateral(items[0][0], items[0][1], 8.0, 8.0, 0.0, 0.1, tune={"graph": -1})
        torch.cuda.synchronize()
        assert np.isfinite(eager.cpu().numpy()).all()
    for rep in range(1, 3):
        for a, b in zip(outs[0], outs[rep]):
            assert np.array_equal(a, b)


def test_single_image_graph_vs_eager(gpu_ctx):
    print("\n[case] single image graph vs eager", flush=True)
    im = S.rect_image(160, 120, 3, K=12, noise=0.05, seed=77)
    a = dev(im)
    g1 = gpu_ctx.joint_bilateral(a, a, 8.0, 8.0, 0.0, 0.1).cpu().numpy()
    g2 = gpu_ctx.joint_bilateral(a, a, 8.0, 8.0, 0.0, 0.1).cpu().numpy()
    e = gpu_ctx.joint_bilateral(a, a, 8.0, 8.0, 0.0, 0.1, tune={"graph": -1}).cpu().numpy()
    assert np.array_equal(g1, g2) and np.array_equal(g1, e)


def test_batch_parts_from_measured_vertices(gpu_ctx):
    """A batch's first call sorts and measures each image's vertex count; a
    repeat of the same images splits the table into per-image part runs in
    proportion to those counts and numbers vertices by hash.  Every call's
    output equals the plain table's bit for bit."""
    import torch
    from imagestack_amd import gtlattice as G
    print("\n[case] batch table parts from measured vertex counts", flush=True)
    imgs = S.batch_images(seed=7, count=12)
    items = []
    for im, ss, sr in imgs:
        a = torch.from_numpy(im).cuda()[:, None]
        items.append((a, a, torch.empty_like(a), ss, sr))
    ctx = G.Context(0)
    try:
        outs = []
        for _ in range(3):
            ctx.joint_bilateral_batch(items)
            outs.append([o.cpu().numpy().copy() for (_, _, o, _, _) in items])
        ctx.joint_bilateral_batch(items, tune={"parts": -1, "neighbours": 1})
        ref = [o.cpu().numpy() for (_, _, o, _, _) in items]
    finally:
        ctx.close()
    for run in outs:
        for a, b in zip(run, ref):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_batch_parts_after_content_change(gpu_ctx):
    """The same image sizes and filters with other content (noise: many more
    vertices) than the measured counts: the split table may overflow; the
    call is then redone on the plain table and the counts measured again.
    Outputs equal the plain table's bit for bit, before and after."""
    import torch
    from imagestack_amd import gtlattice as G
    print("\n[case] batch table parts after the images' content changes", flush=True)
    imgs = S.batch_images(seed=8, count=10)
    rng = np.random.default_rng(5)
    smooth, noisy = [], []
    for im, ss, sr in imgs:
        a = torch.from_numpy(im).cuda()[:, None]
        smooth.append((a, a, torch.empty_like(a), ss, sr))
        n = torch.from_numpy(rng.uniform(size=im.shape).astype(np.float32)).cuda()[:, None]
        noisy.append((n, n, torch.empty_like(n), ss, sr))
    ctx = G.Context(0)
    try:
        retries = []
        for items in (smooth, smooth, noisy, noisy, smooth):
            st = ctx.joint_bilateral_batch(items, stats=True)
            retries.append(st["retries"])
            got = [o.cpu().numpy().copy() for (_, _, o, _, _) in items]
            ctx.joint_bilateral_batch(items, tune={"parts": -1, "neighbours": 1})
            want = [o.cpu().numpy() for (_, _, o, _, _) in items]
            for a, b in zip(got, want):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        ctx.close()
    print("retries per call", retries, flush=True)
    # the first noisy call overflowed the parts sized from the smooth images'
    # counts and was redone; the redo measured the counts again, so the
    # second noisy call (and the smooth one after it) fit at once
    assert retries[2] >= 1
    assert retries[3] == 0 and retries[4] == 0
