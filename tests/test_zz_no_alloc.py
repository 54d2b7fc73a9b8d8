"""gt_joint_bilateral_noalloc (the entry point that allocates nothing)
against gt_joint_bilateral through the C ABI: two calls in a row, the second
smaller, each bit for bit the same as the allocating entry point's output;
then the limits are rejected with GT_EINVAL and the context stays usable."""
import numpy as np
import pytest

from imagestack_amd import gtlattice as G
from imagestack_amd import synth as S

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


CALLS = [  # (W, H, guide channels, sigma_s, sigma_r, storage): the second smaller than the first
    (1024, 1024, 3, 16.0, 0.1, "f32"),
    (300, 200, 3, 4.0, 0.05, "f32"),
    (512, 384, 1, 8.0, 0.1, "bf16"),
]


def test_noalloc_matches_allocating_entry_point():
    """The no-allocation calls run first, back to back: after the first one
    (which takes the scratch array's address on this device) the device's free
    memory does not move.  Then gt_joint_bilateral on a context of its own
    filters the same inputs."""
    import torch
    print("\n[case] no-allocation entry point: calls in a row against gt_joint_bilateral", flush=True)
    inputs, got, free = [], [], []
    na = G.Context(0)
    try:
        for k, (W, H, Cr, ss, sr, storage) in enumerate(CALLS):
            print(f"[case] call {k}: {W}x{H} guide {Cr} channels sigma_s {ss} sigma_r {sr} {storage}", flush=True)
            guide = S.rect_image(W, H, Cr, K=40, noise=0.05, seed=200 + k)
            im = S.uniform_image(W, H, 3, seed=300 + k)
            di, dg = dev(im), dev(guide)
            out = torch.empty_like(di)
            torch.cuda.synchronize()
            before = torch.cuda.mem_get_info()[0]
            _, st = na.joint_bilateral_noalloc(di, dg, ss, ss, 0.0, sr, out, storage=storage, stats=True)
            after = torch.cuda.mem_get_info()[0]
            print("M", st["M"], "capacity", st["capacity"], "free before/after", before, after, flush=True)
            free.append((before, after))
            inputs.append((di, dg, ss, sr, storage))
            got.append(out.cpu().numpy())
    finally:
        na.close()
    for before, after in free[1:]:
        assert before == after
    ref = G.Context(0)
    try:
        for k, ((di, dg, ss, sr, storage), g) in enumerate(zip(inputs, got)):
            want = ref.joint_bilateral(di, dg, ss, ss, 0.0, sr, storage=storage).cpu().numpy()
            assert same_bits(g, want), k
    finally:
        ref.close()


def test_noalloc_limits_are_errors():
    import torch
    print("\n[case] no-allocation entry point: limits", flush=True)
    ctx = G.Context(0)
    try:
        big = torch.zeros((3, 1, 1025, 1024), device="cuda")
        with pytest.raises(G.GTError) as e:
            ctx.joint_bilateral_noalloc(big, big, 8.0, 8.0, 0.0, 0.1, torch.empty_like(big))
        assert e.value.code == 1
        im = torch.rand((3, 1, 64, 64), device="cuda")
        with pytest.raises(G.GTError) as e:   # a zero spatial sigma splits the image
            ctx.joint_bilateral_noalloc(im, im, 0.0, 8.0, 0.0, 0.1, torch.empty_like(im))
        assert e.value.code == 1
        g6 = torch.rand((6, 1, 64, 64), device="cuda")
        with pytest.raises(G.GTError) as e:   # d = 8 lattice
            ctx.joint_bilateral_noalloc(im, g6, 8.0, 8.0, 0.0, 0.1, torch.empty_like(im))
        assert e.value.code == 1
        out = ctx.joint_bilateral_noalloc(im, im, 8.0, 8.0, 0.0, 0.1, torch.empty_like(im))
        want = ctx.joint_bilateral(im, im, 8.0, 8.0, 0.0, 0.1)
        assert same_bits(out.cpu().numpy(), want.cpu().numpy())
    finally:
        ctx.close()
