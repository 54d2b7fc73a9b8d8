"""Non-local means on the GPU (gt_patch_project, gt_nlmeans): the patch
projection bit for bit against the oracle and the reference's own outputs
(golden fixtures), over image borders, ragged tile edges, 3-D patches and
1..16 patch dimensions; the filter against the FP64 oracle within the 1e-5
gate (d = 8 lattice, d = 6 volume, and the grid for 4 position channels);
in-place output; the AUTO cases the GPU does not run."""
import os

import numpy as np
import pytest

from imagestack_amd import gtlattice as G
from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "nlmeans_cases.npz"), allow_pickle=False)


@pytest.mark.parametrize("name", ["rgb", "vol", "grey"])
def test_projection_golden_bits(gpu_ctx, gold, name):
    print(f"\n[case] patch projection golden {name}", flush=True)
    out = gpu_ctx.patch_project(dev(gold[name + "_im"]), dev(gold[name + "_filters"])).cpu().numpy()
    assert np.array_equal(out, gold[name + "_proj"])


@pytest.mark.parametrize("W,H,F,C,K,ft,fh,fw", [
    (37, 29, 1, 3, 6, 1, 7, 7), (16, 16, 1, 1, 1, 1, 1, 1), (100, 3, 1, 2, 5, 1, 3, 5), (5, 70, 2, 3, 9, 1, 9, 3),
    (50, 41, 4, 1, 16, 3, 3, 3), (33, 17, 1, 4, 3, 1, 13, 13), (200, 150, 1, 3, 8, 1, 5, 5)])
def test_projection_vs_oracle_bits(gpu_ctx, oracle_lib, W, H, F, C, K, ft, fh, fw):
    print(f"\n[case] patch projection {W}x{H}x{F}x{C} K={K} filter {fw}x{fh}x{ft}", flush=True)
    rng = np.random.default_rng(W * H + K)
    im = rng.uniform(-1, 1, size=(C, F, H, W)).astype(np.float32)
    filt = rng.normal(size=(K * C, ft, fh, fw)).astype(np.float32)
    out = gpu_ctx.patch_project(dev(im), dev(filt)).cpu().numpy()
    assert np.array_equal(out, oracle_lib.patch_project(im, filt))


@pytest.mark.parametrize("name", ["rgb", "vol", "grey"])
def test_nlmeans_golden(gpu_ctx, oracle_lib, gold, name):
    print(f"\n[case] nlmeans golden {name}", flush=True)
    _, _, ss, sp = (float(v) for v in gold[name + "_params"])
    im, filt = gold[name + "_im"], gold[name + "_filters"]
    out = gpu_ctx.nlmeans(dev(im), dev(filt), ss, sp).cpu().numpy()
    o64 = oracle_lib.nlmeans(im, filt, ss, sp, mode=oracle_lib.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL
    assert rel_err(out, gold[name + "_out"]) < 2 * FP32_REL_TOL


def test_nlmeans_larger_vs_fp64(gpu_ctx, oracle_lib):
    print("\n[case] nlmeans 256x192 RGB, 7x7 patches, 6 dims (d = 8)", flush=True)
    im = S.noisy_rect_image(256, 192, 3, K=24, amp=0.2, seed=45)
    rng = np.random.default_rng(3)
    filt = (rng.normal(size=(18, 1, 7, 7)) / 7).astype(np.float32)
    out, st = gpu_ctx.nlmeans(dev(im), dev(filt), 8.0, 0.3, stats=True)
    o64 = oracle_lib.nlmeans(im, filt, 8.0, 0.3, mode=oracle_lib.F64_LAT)
    assert rel_err(out.cpu().numpy(), o64) < FP32_REL_TOL
    assert st["M"] > 0
    host = gpu_ctx.nlmeans_host(im, filt, 8.0, 0.3)
    assert np.array_equal(host, out.cpu().numpy())


def test_nlmeans_in_place(gpu_ctx, gold):
    print("\n[case] nlmeans in place", flush=True)
    im, filt = gold["rgb_im"], gold["rgb_filters"]
    want = gpu_ctx.nlmeans(dev(im), dev(filt), 16.0, 0.4).cpu().numpy()
    a = dev(im)
    gpu_ctx.nlmeans(a, dev(filt), 16.0, 0.4, out=a)
    assert np.array_equal(a.cpu().numpy(), want)


def test_nlmeans_unsupported(gpu_ctx):
    print("\n[case] nlmeans AUTO cases not on the GPU", flush=True)
    im = dev(S.noisy_rect_image(40, 30, 1, seed=44))
    with pytest.raises(G.GTError) as e:
        gpu_ctx.nlmeans(im, dev(np.ones((9, 1, 3, 3), np.float32)), 4.0, 0.3)
    assert e.value.code == G.GT_EUNSUPPORTED
    with pytest.raises(G.GTError) as e:
        gpu_ctx.nlmeans(im, dev(np.ones((2, 1, 3, 3), np.float32)), 0.3, 0.3)
    assert e.value.code == G.GT_EUNSUPPORTED
    with pytest.raises(G.GTError) as e:
        gpu_ctx.patch_project(im, dev(np.ones((2, 1, 2, 3), np.float32)))
    assert e.value.code == 1
