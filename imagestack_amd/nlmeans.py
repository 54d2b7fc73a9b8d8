"""Non-local means on the GPU behind the reference's NLMeans interface.

NLMeans::apply(image, patchSize, dimensions, spatialSigma, patchSigma)
(src/GaussTransform.cpp:867-874) is PatchPCA (a small host computation over
at most 10 000 random patches, src/Statistics.cpp:1710-1770), a patch
projection (Convolve zero/inner) and a joint bilateral filter over the
projected patches.  `patch_pca` restates PatchPCA on the host with numpy:
the same Gaussian patch weights, sample count, covariance and output layout;
its random patches come from numpy's generator (not the host's rand()) and
its eigenvectors from LAPACK (the reference iterates), so its filters match
the reference's up to sampling noise and eigenvector sign.  The projection
and the filter then run in libgtlattice.so (gt_nlmeans).
"""
from __future__ import annotations

import numpy as np


def patch_side(sigma: float) -> int:
    """((int)(sigma*6+1)) | 1 (src/Statistics.cpp:1712)."""
    return int(np.float32(sigma) * np.float32(6) + np.float32(1)) | 1


def patch_pca(im: np.ndarray, sigma: float, dims: int, seed: int = 0, three_d: bool = False) -> np.ndarray:
    """PatchPCA::apply (or PatchPCA3D::apply when three_d) on a (C, F, H, W)
    float32 image: returns filters (C*dims, side or 1, side, side), plane
    i*C + c holding eigenvector i's weights for channel c.  PatchPCA3D's own
    choices are kept: 1000 samples instead of 10 000, and each tap reads frame
    t + dy, not t + dt (src/Statistics.cpp:1814, 1826)."""
    im = np.asarray(im, np.float32)
    Cc, F, H, W = im.shape
    side = patch_side(sigma)
    r = side // 2
    i = np.arange(side, dtype=np.float32) - np.float32(r)
    mask = np.exp(-(i * i) / np.float32(2 * sigma * sigma)).astype(np.float32)
    mask = (mask / mask.sum(dtype=np.float32)).astype(np.float32)
    if W <= 2 * r or H <= 2 * r or (three_d and F <= 2 * r):
        raise ValueError("image smaller than one patch")
    count = min(1000 if three_d else 10000, W * H * F)
    rng = np.random.default_rng(seed)
    xs = rng.integers(r, W - r, count)
    ys = rng.integers(r, H - r, count)
    ts = rng.integers(r, F - r, count) if three_d else rng.integers(0, F, count)
    d = np.arange(-r, r + 1)
    if three_d:
        w = (mask[:, None, None] * mask[None, :, None] * mask[None, None, :])            # (dt, dy, dx)
        tt = np.broadcast_to(ts[:, None, None, None] + d[None, None, :, None], (count, side, side, side))
        pt = im[:, tt, ys[:, None, None, None] + d[None, None, :, None],
                xs[:, None, None, None] + d[None, None, None, :]]                         # (C, n, dt, dy, dx)
        vec = (pt * w[None, None]).transpose(1, 2, 3, 4, 0).reshape(count, -1)          # order dt, dy, dx, c
    else:
        w = mask[:, None] * mask[None, :]                                                # (dy, dx)
        pt = im[:, ts[:, None, None], ys[:, None, None] + d[:, None], xs[:, None, None] + d[None, :]]
        vec = (pt * w[None, None]).transpose(1, 2, 3, 0).reshape(count, -1)             # order dy, dx, c
    vec = vec.astype(np.float64)
    mean = vec.sum(0)
    cov = (vec.T @ vec - np.outer(mean, mean) / count) / count     # eigenvectors.h compute(): mean removed
    evals, evecs = np.linalg.eigh(cov)
    top = evecs[:, np.argsort(evals)[::-1][:dims]]                  # (D, dims), largest first
    ft = side if three_d else 1
    # vec index j = ((dt*side + dy)*side + dx)*C + c -> filters(x, y, t, i*C + c)
    f = top.T.reshape(dims, ft, side, side, Cc).transpose(0, 4, 1, 2, 3)
    return np.ascontiguousarray(f.reshape(dims * Cc, ft, side, side), dtype=np.float32)


def nlmeans(ctx, im, sigma: float, dims: int, spatial_sigma: float, patch_sigma: float, seed: int = 0,
            three_d: bool = False, out=None):
    """NLMeans::apply (three_d: NLMeans3D::apply) on a (C, F, H, W) device
    tensor through a gtlattice.Context: host PCA, then gt_nlmeans."""
    import torch
    filters = patch_pca(im.detach().cpu().numpy(), sigma, dims, seed=seed, three_d=three_d)
    return ctx.nlmeans(im, torch.from_numpy(filters).to(im.device), spatial_sigma, patch_sigma, out=out)
