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
