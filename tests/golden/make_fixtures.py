"""Writes the golden fixtures in tests/golden/ from the reference itself,
compiled from its own sources by oracle/Makefile (oracle/_ref/libref_strict.so,
strict IEEE flags).  Run in a container that has the reference sources:

    make -C oracle all && python tests/golden/make_fixtures.py

Fixtures are data only: seeded inputs, reference outputs, sorted lattice keys
and digests.  Small cases hold full arrays; the 512^2 and 2048^2 cases hold
M and SHA-256 digests of the sorted key set and of the output.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from imagestack_amd import synth as S  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def sorted_ref_lattice(pos, val, inv, blur):
    keys, vals = O.ref_lattice(pos, val, inv, blur=blur)
    order = O.sort_keys(keys)
    return keys[order], vals[order]


def gauss_testcase():
    """GaussTransform::test shape (src/GaussTransform.cpp:39-77): 56x34x2,
    3-D positions U[0,1], values [U[0,1], 1], sigmas (0.2, 0.3, 0.4), slice
    positions distinct from splat positions (non-replay slice)."""
    s = S.Stream(11)
    W, H, F = 56, 34, 2
    n = W * H * F
    splat = s.uniform(3 * n).reshape(3, F, H, W).astype(np.float32)
    slc = s.uniform(3 * n).reshape(3, F, H, W).astype(np.float32)
    vals = np.stack([s.uniform(n), np.ones(n)]).reshape(2, F, H, W).astype(np.float32)
    sig = np.array([0.2, 0.3, 0.4], np.float32)
    out = O.ref_gauss_transform(splat, vals, sig, slice_pos=slc)
    out_replay = O.ref_gauss_transform(splat, vals, sig)
    inv = (np.float32(1) / sig).astype(np.float32)
    keys, splat_v = sorted_ref_lattice(splat, vals, inv, blur=False)
    _, blur_v = sorted_ref_lattice(splat, vals, inv, blur=True)
    np.savez_compressed(os.path.join(HERE, "gausstransform_56x34x2.npz"), splat=splat, slice=slc, values=vals,
                        sigmas=sig, out=out, out_replay=out_replay, keys=keys, splat_vals=splat_v,
                        blur_vals=blur_v)


def bilateral_small():
    """64x48 RGB rectangles + N(0,0.05), -bilateral 0.1 8: d=5, vd=4."""
    im = S.rect_image(64, 48, 3, K=8, noise=0.05, seed=64)
    out = O.ref_joint_bilateral(im, im, 8.0, 8.0, 0.0, 0.1)
    pos = S.bilateral_positions(im, 8.0, 0.1)
    val = S.homogeneous_values(im)
    inv = np.ones(5, np.float32)
    raw = O.ref_gauss_transform(pos, val, np.ones(5, np.float32))
    keys, splat_v = sorted_ref_lattice(pos, val, inv, blur=False)
    _, blur_v = sorted_ref_lattice(pos, val, inv, blur=True)
    np.savez_compressed(os.path.join(HERE, "bilateral_64x48.npz"), im=im, out=out, raw=raw, keys=keys,
                        splat_vals=splat_v, blur_vals=blur_v)


def jointbilateral_testcase():
    """JointBilateral::test shape (src/GaussTransform.cpp:442-471): 60x50x3
    frames, RGB image and guide U[0,1], sigmas (3, 4, 1), colour sigma 0.25:
    a d=6 lattice with a temporal axis."""
    s = S.Stream(12)
    im = s.uniform(3 * 3 * 50 * 60).reshape(3, 3, 50, 60).astype(np.float32)
    ref = s.uniform(3 * 3 * 50 * 60).reshape(3, 3, 50, 60).astype(np.float32)
    out = O.ref_joint_bilateral(im, ref, 3.0, 4.0, 1.0, 0.25)
    np.savez_compressed(os.path.join(HERE, "jointbilateral_60x50x3.npz"), im=im, ref=ref, out=out)


def zero_sigma_case():
    """sigma_y = 0: rows filtered independently (src/GaussTransform.cpp:524-530)."""
    im = S.rect_image(40, 12, 3, K=6, noise=0.05, seed=21)
    out = O.ref_joint_bilateral(im, im, 4.0, 0.0, 0.0, 0.2)
    np.savez_compressed(os.path.join(HERE, "bilateral_rows_40x12.npz"), im=im, out=out)


def highdim_small():
    """d=16 Gauss transform at 32x32 (14 N(0,1) features + x/12, y/12)."""
    pos = S.highdim_positions(32, 32, nfeat=14, sigma_s=12.0, seed=3)
    s = S.Stream(4)
    val = s.uniform(4 * 32 * 32).reshape(4, 1, 32, 32).astype(np.float32)
    out = O.ref_gauss_transform(pos, val, np.ones(16, np.float32))
    keys, _ = sorted_ref_lattice(pos, val, np.ones(16, np.float32), blur=False)
    np.savez_compressed(os.path.join(HERE, "highdim_d16_32x32.npz"), pos=pos, values=val, out=out, keys=keys)


def grid_cases():
    """The GRID method (src/DenseGrid.h via src/GaussTransform.cpp:227-297):
    a grey bilateral (d=3), an RGB bilateral (d=5) and a Gauss transform with
    distinct slice positions (d=3, vd=2), reference outputs."""
    grey = S.rect_image(60, 44, 1, K=6, noise=0.05, seed=31)
    out_grey = O.ref_joint_bilateral(grey, grey, 4.0, 4.0, 0.0, 0.1, grid=True)
    rgb = S.rect_image(48, 36, 3, K=6, noise=0.05, seed=32)
    out_rgb = O.ref_joint_bilateral(rgb, rgb, 3.0, 3.0, 0.0, 0.2, grid=True)
    s = S.Stream(13)
    pos = s.normal(3 * 30 * 20).reshape(3, 1, 20, 30).astype(np.float32)
    slc = s.normal(3 * 30 * 20).reshape(3, 1, 20, 30).astype(np.float32)
    vals = np.stack([s.uniform(600), np.ones(600)]).reshape(2, 1, 20, 30).astype(np.float32)
    sig = np.array([0.5, 0.7, 1.2], np.float32)
    out_gt = O.ref_grid_transform(pos, vals, sig, slice_pos=slc)
    out_gt_replay = O.ref_grid_transform(pos, vals, sig)
    np.savez_compressed(os.path.join(HERE, "grid_cases.npz"), grey=grey, out_grey=out_grey, rgb=rgb,
                        out_rgb=out_rgb, pos=pos, slice=slc, values=vals, sigmas=sig, out_gt=out_gt,
                        out_gt_replay=out_gt_replay)


def nlmeans_cases():
    """NLMeans::apply (src/GaussTransform.cpp:867-874) and NLMeans3D::apply
    (:934-941) with the reference's own PatchPCA after srand(seed): the
    filters, the patch projection (Convolve zero/inner) and the filtered image.
    Cases: RGB 64x48 (7x7 patches, 6 dims: a d=8 lattice, NLMeans::test's
    sigmas 16 and 0.4); RGB volume 24x20x8 (5x5x5 patches, 4 dims: d=6); grey
    48x40 with 2 dims (4 position channels: AUTO picks the grid)."""
    rec = {}
    cases = [("rgb", S.noisy_rect_image(64, 48, 3, K=10, amp=0.2, seed=41), 1.0, 6, 16.0, 0.4, False, 7),
             ("vol", S.noisy_rect_image(24, 20, 3, K=6, amp=0.2, seed=42, F=8), 0.5, 4, 4.0, 0.3, True, 8),
             ("grey", S.noisy_rect_image(48, 40, 1, K=8, amp=0.2, seed=43), 1.0, 2, 6.0, 0.15, False, 9)]
    for name, im, sigma, dims, ss, sp, three_d, seed in cases:
        filt = O.ref_patch_pca(im, sigma, dims, seed=seed, three_d=three_d)
        rec[name + "_im"] = im
        rec[name + "_filters"] = filt
        rec[name + "_proj"] = O.ref_convolve_inner(im, filt)
        rec[name + "_out"] = O.ref_nlmeans(im, sigma, dims, ss, sp, seed=seed, three_d=three_d)
        rec[name + "_params"] = np.array([sigma, dims, ss, sp], np.float32)
    np.savez_compressed(os.path.join(HERE, "nlmeans_cases.npz"), **rec)


TMP_TYPES = ["float32", "float64", "uint8", "int8", "uint16", "int16", "uint32", "int32", "uint64", "int64"]


def tmp_cases():
    """The .tmp format as the reference writes and reads it (FileTMP::save /
    load, src/FileTMP.cpp:75-233), through the reference CLI: a 7x5x2x3 float
    image with fractional, negative and mid-range samples saved as every type
    (tmp/as_<type>.tmp), and each of those loaded back and saved as float32
    (tmp/back_<type>.tmp)."""
    import subprocess
    import tempfile
    cli = os.path.join(ROOT, "oracle", "_ref", "ImageStack")
    d = os.path.join(HERE, "tmp")
    os.makedirs(d, exist_ok=True)
    s = S.Stream(17)
    im = (s.uniform(3 * 2 * 5 * 7) * 220.0 - 100.0).reshape(3, 2, 5, 7).astype(np.float32)
    im.reshape(-1)[:6] = [0.0, -0.0, 0.5, -0.5, 99.99, -99.99]
    with tempfile.TemporaryDirectory() as tdir:
        src = os.path.join(tdir, "in.tmp")
        with open(src, "wb") as f:
            np.array([7, 5, 2, 3, 0], np.int32).tofile(f)
            im.tofile(f)
        for t in TMP_TYPES:
            out = os.path.join(d, f"as_{t}.tmp")
            subprocess.run([cli, "-load", src, "-save", out, t], check=True, capture_output=True)
            subprocess.run([cli, "-load", out, "-save", os.path.join(d, f"back_{t}.tmp")], check=True,
                           capture_output=True)
        with open(os.path.join(d, "source.tmp"), "wb") as f:
            np.array([7, 5, 2, 3, 0], np.int32).tofile(f)
            im.tofile(f)


def digests():
    """The 512^2 bilateral, the 2048^2 joint bilateral and the 1024^2 d = 16
    transform: M and digests (key sets; outputs where the reference's FP32
    bits are the parity target of the oracle)."""
    rec = {}
    im, ref, ss, sr = S.bilateral_512()
    pos = S.bilateral_positions(ref, ss, sr)
    val = S.homogeneous_values(im)
    keys, _ = sorted_ref_lattice(pos, val, np.ones(5, np.float32), blur=False)
    out = O.ref_joint_bilateral(im, ref, ss, ss, 0.0, sr)
    rec["bilateral_512"] = {"M": int(keys.shape[0]), "keys_sha": digest(keys), "out_sha": digest(out),
                                 "input_sha": digest(im)}
    im, ref, ss, sr = S.joint_bilateral_2048()
    pos = S.bilateral_positions(ref, ss, sr)
    val = S.homogeneous_values(im)
    keys, _ = sorted_ref_lattice(pos, val, np.ones(5, np.float32), blur=False)
    out = O.ref_joint_bilateral(im, ref, ss, ss, 0.0, sr)
    rec["joint_bilateral_2048"] = {"M": int(keys.shape[0]), "keys_sha": digest(keys), "out_sha": digest(out),
                                       "guide_sha": digest(ref), "image_sha": digest(im)}
    pos, val = S.highdim_1024()
    keys, _ = sorted_ref_lattice(pos, val, np.ones(16, np.float32), blur=False)
    rec["highdim_1024"] = {"M": int(keys.shape[0]), "keys_sha": digest(keys), "pos_sha": digest(pos),
                           "values_sha": digest(val)}
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if not O.ref_available("strict"):
        sys.exit("oracle/_ref/libref_strict.so is missing: run `make -C oracle all` with the reference present")
    gauss_testcase()
    bilateral_small()
    jointbilateral_testcase()
    zero_sigma_case()
    highdim_small()
    grid_cases()
    nlmeans_cases()
    tmp_cases()
    digests()
    print("fixtures written to", HERE)
