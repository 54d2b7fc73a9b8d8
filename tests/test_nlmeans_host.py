"""The host PatchPCA restatement (imagestack_amd/nlmeans.py) against the
reference's own PatchPCA / PatchPCA3D filters (golden fixtures).  Sampling
differs (numpy's generator, not rand()), so the check is on what sampling
does not move: the span of the two leading eigenvectors, the patch side, the
plane layout (a transposed or channel-swapped layout fails the span check),
and unit-norm eigenvectors."""
import os

import numpy as np
import pytest

from imagestack_amd import nlmeans as NL

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def span_overlap(a, b, dims, C, k=2):
    a = a.reshape(dims, C, -1).transpose(0, 2, 1).reshape(dims, -1)[:k].T
    b = b.reshape(dims, C, -1).transpose(0, 2, 1).reshape(dims, -1)[:k].T
    qa, _ = np.linalg.qr(a)
    qb, _ = np.linalg.qr(b)
    return np.linalg.svd(qa.T @ qb)[1].min()


@pytest.mark.parametrize("name,three_d", [("rgb", False), ("vol", True)])
def test_patch_pca_matches_reference_subspace(name, three_d):
    g = np.load(os.path.join(GOLD, "nlmeans_cases.npz"), allow_pickle=False)
    sigma, dims = float(g[name + "_params"][0]), int(g[name + "_params"][1])
    im, ref = g[name + "_im"], g[name + "_filters"]
    f = NL.patch_pca(im, sigma, dims, seed=3, three_d=three_d)
    assert f.shape == ref.shape
    assert span_overlap(f, ref, dims, im.shape[0]) > 0.99
    v = f.reshape(dims, im.shape[0], -1)
    np.testing.assert_allclose((v * v).sum((1, 2)), 1.0, rtol=1e-5)


def test_patch_side():
    assert [NL.patch_side(s) for s in (0.5, 1.0, 2.0, 0.1)] == [5, 7, 13, 1]
