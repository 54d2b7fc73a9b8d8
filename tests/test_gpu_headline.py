"""GPU parity at the full sizes of the measured workloads.

  * 2048^2 joint bilateral (guide: 256 rectangles + N(0, 0.03); image
    U[0, 1]; sigma_s 16, sigma_r 0.1; d = 5, vd = 4): the GPU's sorted key
    set must be the reference's (M and SHA-256 written from the reference,
    tests/golden/digests.json), the neighbour table bit-exact against the
    oracle, and the production call's output within 1e-5 of the FP64 oracle
    and of the reference's own FP32 output (whose bits the F32 oracle
    reproduces: its digest is the reference's).
  * 1024^2 d = 16 Gauss transform (14 N(0,1) features + x/12, y/12): the
    GPU's sorted key set of 17.8 M vertices must be the reference's; its
    output is checked against the FP64 oracle on a 96 x 96 crop (the oracle
    takes minutes over the full frame).
"""
import json
import os

import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, digest, rel_err, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def record(name):
    return json.load(open(os.path.join(GOLD, "digests.json")))[name]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def planar(a):
    a = np.ascontiguousarray(a, np.float32)
    return a.reshape(a.shape[0], -1)


def case(name):
    print(f"\n[case] {name}", flush=True)


@pytest.fixture(scope="module")
def jb2048():
    im, guide, ss, sr = S.joint_bilateral_2048()
    rec = record("joint_bilateral_2048")
    assert digest(guide) == rec["guide_sha"] and digest(im) == rec["image_sha"]
    return im, guide, ss, sr, rec


def test_joint_bilateral_2048_key_set_and_neighbours(gpu_ctx, oracle_lib, jb2048):
    O = oracle_lib
    im, guide, ss, sr, rec = jb2048
    case("joint bilateral 2048x2048 lattice export")
    pos = planar(S.bilateral_positions(guide, ss, sr))
    val = planar(S.homogeneous_values(im))
    inv = np.ones(5, np.float32)
    g = gpu_ctx.build_lattice(dev(pos), dev(val), inv, row_w=2048)
    print("M gpu", g["M"], "reference", rec["M"], flush=True)
    assert g["M"] == rec["M"]
    assert digest(g["keys"]) == rec["keys_sha"]
    o = O.lattice(pos, val, inv, mode=O.F64_LAT)
    assert np.array_equal(g["keys"], o["keys"])
    assert np.array_equal(g["nbr"], o["nbr"])
    assert rel_l2(g["splat"], o["splat"]) < 1e-6
    assert rel_l2(g["blur"], o["blur"]) < 1e-6


def test_joint_bilateral_2048_output(gpu_ctx, oracle_lib, jb2048):
    """The production call (device-built positions, default tuning, cached
    graph) against the FP64 oracle and the reference's FP32 output."""
    O = oracle_lib
    im, guide, ss, sr, rec = jb2048
    case("joint bilateral 2048x2048 output, default tuning")
    di, dg = dev(im), dev(guide)
    out1, st = gpu_ctx.joint_bilateral(di, dg, ss, ss, 0.0, sr, stats=True)
    out1 = out1.cpu().numpy()
    out2 = gpu_ctx.joint_bilateral(di, dg, ss, ss, 0.0, sr).cpu().numpy()   # graph replay
    print("M", st["M"], "retries", st["retries"], flush=True)
    assert st["M"] == rec["M"]
    assert np.array_equal(out1, out2)
    o32 = O.joint_bilateral(im, guide, ss, ss, 0.0, sr)
    assert digest(o32) == rec["out_sha"]   # the reference's own FP32 output
    o64 = O.joint_bilateral(im, guide, ss, ss, 0.0, sr, mode=O.F64_LAT)
    e64, e32 = rel_err(out1, o64), rel_err(out1, o32)
    print("err vs FP64", e64, "vs reference FP32", e32, flush=True)
    assert e64 < FP32_REL_TOL
    assert e32 < FP32_REL_TOL


@pytest.mark.parametrize("storage,l2_bar", [("bf16", 5e-3), ("f16", 1e-3)])
def test_joint_bilateral_2048_narrow_storage(gpu_ctx, oracle_lib, jb2048, storage, l2_bar):
    O = oracle_lib
    im, guide, ss, sr, rec = jb2048
    case(f"joint bilateral 2048x2048 output, {storage} storage")
    out = gpu_ctx.joint_bilateral(dev(im), dev(guide), ss, ss, 0.0, sr, storage=storage).cpu().numpy()
    o64 = O.joint_bilateral(im, guide, ss, ss, 0.0, sr, mode=O.F64_LAT)
    e = rel_l2(out, o64)
    print(storage, "rel L2 vs FP64", e, flush=True)
    assert e < l2_bar


def test_highdim_1024_key_set(gpu_ctx):
    rec = record("highdim_1024")
    pos, val = S.highdim_1024()
    assert digest(pos) == rec["pos_sha"] and digest(val) == rec["values_sha"]
    case("d=16 gauss transform 1024x1024 lattice keys")
    g = gpu_ctx.build_lattice(dev(planar(pos)), dev(planar(val)), np.ones(16, np.float32), row_w=1024,
                              parts=("keys",))
    print("M gpu", g["M"], "reference", rec["M"], flush=True)
    assert g["M"] == rec["M"]
    assert digest(g["keys"]) == rec["keys_sha"]


def test_highdim_1024_crop_output(gpu_ctx, oracle_lib):
    O = oracle_lib
    pos, val = S.highdim_1024()
    c = 96
    pc = planar(pos[:, :, :c, :c])
    vc = planar(val[:, :, :c, :c])
    inv = np.ones(16, np.float32)
    case("d=16 gauss transform 96x96 crop output")
    out = gpu_ctx.gauss_transform(dev(pc), dev(vc), inv, row_w=c).cpu().numpy()
    o64, M = O.gauss_transform(pc, vc, inv, mode=O.F64_LAT)
    e = rel_err(out, o64.reshape(out.shape))
    print("M", M, "err vs FP64", e, flush=True)
    assert e < FP32_REL_TOL
