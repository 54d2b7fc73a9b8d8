"""Lattice dimensions 7..10 (JointBilateral's AUTO range above the plain
bilateral: non-local means, multi-channel guides) through the LDS-tiled
splat and slice: key set, neighbour table and splat/blur values against the
oracle, the Gauss transform against the FP64 oracle (replay and distinct
slice), packed and wide keys, with the tiles on and off."""
import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err, rel_l2

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def positions(d, W, H, spread, seed):
    """d-2 smooth-ish feature channels (rectangles + noise) scaled by
    `spread`, then x/8 and y/8: lattices with shared vertices, like a guide."""
    im = S.rect_image(W, H, d - 2, K=12, noise=0.05, seed=seed)
    x = np.tile(np.arange(W, dtype=np.float32) / 8, (H, 1))
    y = np.tile((np.arange(H, dtype=np.float32) / 8)[:, None], (1, W))
    return np.concatenate([im.reshape(d - 2, -1) * spread, x.reshape(1, -1), y.reshape(1, -1)]).astype(np.float32)


@pytest.mark.parametrize("d,vd,spread", [(7, 4, 4.0), (8, 4, 3.0), (9, 2, 5.0), (10, 4, 2.0), (8, 4, 900.0)])
@pytest.mark.parametrize("dedup", [0, -1])
def test_mid_d_lattice_and_transform(gpu_ctx, oracle_lib, d, vd, spread, dedup):
    O = oracle_lib
    W, H = 96, 64
    print(f"\n[case] d={d} vd={vd} spread={spread} dedup={dedup}", flush=True)
    pos = positions(d, W, H, spread, seed=d * 7 + vd)
    rng = np.random.default_rng(d)
    val = np.concatenate([rng.uniform(size=(vd - 1, W * H)), np.ones((1, W * H))]).astype(np.float32)
    inv = np.ones(d, np.float32)
    tune = {"dedup": dedup}
    g = gpu_ctx.build_lattice(dev(pos), dev(val), inv, row_w=W, tune=tune)
    o = O.lattice(pos, val, inv, mode=O.F64_LAT)
    print("M gpu", g["M"], "oracle", o["M"], flush=True)
    assert g["M"] == o["M"]
    assert np.array_equal(g["keys"], o["keys"])
    assert np.array_equal(g["nbr"], o["nbr"])
    assert rel_l2(g["splat"], o["splat"]) < 1e-6
    assert rel_l2(g["blur"], o["blur"]) < 1e-6
    out = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune=tune).cpu().numpy()
    o64, _ = O.gauss_transform(pos, val, inv, mode=O.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL
    slc = (pos + rng.normal(scale=0.3, size=pos.shape)).astype(np.float32)
    out = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, slice_pos=dev(slc), tune=tune).cpu().numpy()
    o64, _ = O.gauss_transform(pos, val, inv, slice_pos=slc, mode=O.F64_LAT)
    assert rel_err(out, o64) < FP32_REL_TOL


@pytest.mark.parametrize("d", [5, 8, 12, 16])
def test_hash_and_merge_neighbours_agree(gpu_ctx, oracle_lib, d):
    """Vertex ids in compaction order with hash-lookup neighbours (no sort)
    give the same output, bit for bit, as sorted ids with merged neighbours."""
    print(f"\n[case] neighbours by hash vs merge, d={d}", flush=True)
    W, H = 80, 60
    pos = positions(d, W, H, 3.0, seed=d)
    rng = np.random.default_rng(d + 1)
    val = np.concatenate([rng.uniform(size=(3, W * H)), np.ones((1, W * H))]).astype(np.float32)
    inv = np.ones(d, np.float32)
    outs = [gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune={"neighbours": m}).cpu().numpy()
            for m in (1, 2, 0)]
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    o64, _ = oracle_lib.gauss_transform(pos, val, inv, mode=oracle_lib.F64_LAT)
    assert rel_err(outs[1], o64) < FP32_REL_TOL


def test_sparse_one_end_neighbours_with_fill_kernel(gpu_ctx, oracle_lib):
    """A sparse d = 8 lattice large enough (about 9 vertices per pixel, 1.1 M
    vertices) that, once its vertex density is known, the hash-id form looks
    neighbours up from one end of each edge and fills the entries with a
    kernel of its own: the same output, bit for bit, as sorted ids with
    merged neighbours, and within the FP64 gate."""
    print("\n[case] sparse d=8 352x352, one-end neighbours + fill kernel vs merge", flush=True)
    W = H = 352
    pos = S.highdim_positions(W, H, nfeat=6, sigma_s=2.0, seed=11).reshape(8, -1)
    rng = np.random.default_rng(12)
    val = np.concatenate([rng.uniform(size=(3, W * H)), np.ones((1, W * H))]).astype(np.float32)
    inv = np.ones(8, np.float32)
    first, st = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, stats=True)   # learns the density
    assert st["M"] * 8 * 9 > (64 << 20) // 1.05   # large enough for the fill kernel
    hashed = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune={"neighbours": 2}).cpu().numpy()
    merged = gpu_ctx.gauss_transform(dev(pos), dev(val), inv, row_w=W, tune={"neighbours": 1}).cpu().numpy()
    assert np.array_equal(hashed, merged)
    assert np.array_equal(hashed, first.cpu().numpy())
    o64, _ = oracle_lib.gauss_transform(pos, val, inv, mode=oracle_lib.F64_LAT)
    assert rel_err(hashed, o64) < FP32_REL_TOL
