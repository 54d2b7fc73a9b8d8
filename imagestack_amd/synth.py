"""Seeded synthetic inputs for the bilateral / Gauss-transform workloads.

Counter-based SplitMix64 streams (vectorised in numpy), so every size of every
workload is reproducible bit for bit on any host.  Images are dense planar
float32 arrays of shape (C, F, H, W), the layout ImageStack uses for a dense
image (src/Image.h:34-37: element (x,y,t,c) at c*cstride + t*tstride + y*W + x).

Workloads (all seeded):
  * ``rect_image``: K axis-aligned constant-colour rectangles with U[0,1]
    colours painted in order over a zero background, plus iid N(0, sigma)
    noise per channel.
  * ``bilateral_positions``: positions in ImageStack's JointBilateral order
    (colour channels, x, y) with the reciprocal-multiply rule
    (src/GaussTransform.cpp:654-666, src/Expr.h:2439-2458).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """Counter-based SplitMix64: value i is mix(seed_state + (i+1)*gamma)."""

    def __init__(self, seed: int, stream: int = 0):
        self.state = np.uint64((seed * 0x632BE59BD9B4E019 + stream * 0x2545F4914F6CDD1D) & (2**64 - 1))
        self.counter = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(self.counter + 1, self.counter + n + 1, dtype=np.uint64)
            z = self.state + idx * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.counter += n
        return z

    def uniform(self, n: int) -> np.ndarray:
        """U[0,1) doubles from the top 53 bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """N(0,1) doubles by Box-Muller."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        th = 2.0 * np.pi * u2
        return np.concatenate([r * np.cos(th), r * np.sin(th)])[:n]

    def integers(self, lo: int, hi: int, n: int) -> np.ndarray:
        """Uniform integers in [lo, hi)."""
        return lo + np.floor(self.uniform(n) * (hi - lo)).astype(np.int64)


def rect_image(W: int, H: int, C: int = 3, K: int = 64, noise: float = 0.05,
               seed: int = 1, F: int = 1) -> np.ndarray:
    """Piecewise-constant RGB image: K rectangles + N(0, noise). (C,F,H,W) f32."""
    s = Stream(seed, 1)
    img = np.zeros((C, F, H, W), dtype=np.float64)
    xs = s.integers(0, W, 2 * K).reshape(K, 2)
    ys = s.integers(0, H, 2 * K).reshape(K, 2)
    cols = s.uniform(C * K).reshape(K, C)
    for k in range(K):
        x0, x1 = sorted(xs[k])
        y0, y1 = sorted(ys[k])
        img[:, :, y0:y1 + 1, x0:x1 + 1] = cols[k][:, None, None, None]
    if noise > 0:
        img += noise * s.normal(img.size).reshape(img.shape)
    return img.astype(np.float32)


def noisy_rect_image(W: int, H: int, C: int = 3, K: int = 16, amp: float = 0.2, seed: int = 41,
                     F: int = 1) -> np.ndarray:
    """Denoising input: K rectangles plus iid U[-amp, amp] noise, the kind of
    noise NLMeans::test adds (src/GaussTransform.cpp:824-838). (C,F,H,W) f32."""
    clean = rect_image(W, H, C, K=K, noise=0.0, seed=seed, F=F).astype(np.float64)
    u = Stream(seed, 3).uniform(clean.size).reshape(clean.shape)
    return (clean + amp * (2.0 * u - 1.0)).astype(np.float32)


def uniform_image(W: int, H: int, C: int = 3, seed: int = 2, F: int = 1) -> np.ndarray:
    s = Stream(seed, 2)
    return s.uniform(C * F * H * W).reshape(C, F, H, W).astype(np.float32)


def bilateral_positions(ref: np.ndarray, sigma_s: float, sigma_r: float,
                        sigma_t: float = 0.0) -> np.ndarray:
    """Positions (Cr + 2 [+1], F, H, W) as JointBilateral::apply builds them.

    ref * (1.0f/sigma_r) per colour channel, then float(x) * (1.0f/sigma_s),
    float(y) * (1.0f/sigma_s) (src/GaussTransform.cpp:654-666).  Spatial
    channels are included only when that axis is filtered, as in the reference
    (filterX = W > 1 and sigma_s < 10 W, :541-543).
    """
    C, F, H, W = ref.shape
    f32 = np.float32
    chans = [ref[c] * (f32(1.0) / f32(sigma_r)) for c in range(C)]
    if W > 1 and sigma_s != 0 and sigma_s < 10 * W:
        xx = np.broadcast_to(np.arange(W, dtype=f32)[None, None, :], (F, H, W))
        chans.append(xx * (f32(1.0) / f32(sigma_s)))
    if H > 1 and sigma_s != 0 and sigma_s < 10 * H:
        yy = np.broadcast_to(np.arange(H, dtype=f32)[None, :, None], (F, H, W))
        chans.append(yy * (f32(1.0) / f32(sigma_s)))
    if F > 1 and sigma_t != 0 and sigma_t < 10 * F:
        tt = np.broadcast_to(np.arange(F, dtype=f32)[:, None, None], (F, H, W))
        chans.append(tt * (f32(1.0) / f32(sigma_t)))
    return np.stack([np.ascontiguousarray(c, dtype=f32) for c in chans])


def homogeneous_values(im: np.ndarray) -> np.ndarray:
    """[im, 1] as JointBilateral builds the value image (:647-651)."""
    C, F, H, W = im.shape
    return np.concatenate([im, np.ones((1, F, H, W), np.float32)]).astype(np.float32)


def highdim_positions(W: int, H: int, nfeat: int = 14, sigma_s: float = 12.0,
                      seed: int = 3) -> np.ndarray:
    """d = nfeat + 2 positions: N(0,1) features (unit lattice spacing) then
    x/sigma_s, y/sigma_s.  (nfeat+2, 1, H, W) f32."""
    s = Stream(seed, 3)
    f32 = np.float32
    feats = s.normal(nfeat * H * W).reshape(nfeat, 1, H, W).astype(f32)
    xx = np.broadcast_to(np.arange(W, dtype=f32)[None, None, :], (1, H, W)) * (f32(1.0) / f32(sigma_s))
    yy = np.broadcast_to(np.arange(H, dtype=f32)[None, :, None], (1, H, W)) * (f32(1.0) / f32(sigma_s))
    return np.concatenate([feats, xx[None], yy[None]]).astype(f32)


def highdim_1024(side: int = 1024):
    """d = 16 Gauss transform (non-local-means style) at side x side: 14
    N(0,1) features + x/12, y/12 (seed 3), values 4 x U[0,1] (seed 33).
    Returns (positions (16, 1, side, side), values (4, 1, side, side))."""
    pos = highdim_positions(side, side, nfeat=14, sigma_s=12.0, seed=3)
    vals = uniform_image(side, side, 4, seed=33)
    return pos, vals


def bilateral_512(seed: int = 1):
    """512x512 RGB bilateral: 64 rectangles + N(0,0.05); sigma_s=8, sigma_r=0.1."""
    im = rect_image(512, 512, 3, K=64, noise=0.05, seed=seed)
    return im, im, 8.0, 0.1


def joint_bilateral_2048(seed: int = 1, side: int = 2048):
    """2048^2 joint bilateral: guide 256 rectangles + N(0,0.03), image U[0,1];
    sigma_s=16, sigma_r=0.1."""
    guide = rect_image(side, side, 3, K=256, noise=0.03, seed=seed)
    im = uniform_image(side, side, 3, seed=seed + 100)
    return im, guide, 16.0, 0.1


def batch_images(seed: int = 4, count: int = 64):
    """Batch of `count` RGB bilateral images: sides log-uniform in [256, 1536],
    drawn independently per axis; content as bilateral_512 (64 rectangles +
    N(0, 0.05)); sigma_s from {4, 8, 16} and sigma_r from {0.05, 0.1, 0.2}
    per image.  Returns [(im (3, H, W), sigma_s, sigma_r)]."""
    s = Stream(seed, 11)
    lo, hi = np.log(256.0), np.log(1536.0)
    sides = np.rint(np.exp(lo + s.uniform(2 * count) * (hi - lo))).astype(np.int64).reshape(count, 2)
    ss = np.array([4.0, 8.0, 16.0])[s.integers(0, 3, count)]
    sr = np.array([0.05, 0.1, 0.2])[s.integers(0, 3, count)]
    out = []
    for i in range(count):
        W, H = int(sides[i, 0]), int(sides[i, 1])
        im = rect_image(W, H, 3, K=64, noise=0.05, seed=seed * 1000 + i)[:, 0]
        out.append((np.ascontiguousarray(im), float(ss[i]), float(sr[i])))
    return out


def serving_schedule(n: int, qps: float, seed: int = 5, pool_side: int = 4096):
    """Serving workload: Poisson arrivals at `qps`; square sides log-uniform
    in [128, 4096]; 60% grey (C=1, d=3, vd=2) and 40% RGB (C=3, d=5, vd=4);
    sigma_s from {4, 16, 32}, sigma_r from {0.05, 0.2}.  Each request is a crop
    at a uniform offset of a pool image.  Returns dict of arrays."""
    s = Stream(seed, 13)
    arrival = np.cumsum(-np.log1p(-s.uniform(n)) / qps * 1000.0)
    arrival -= arrival[0]
    lo, hi = np.log(128.0), np.log(4096.0)
    side = np.minimum(np.rint(np.exp(lo + s.uniform(n) * (hi - lo))).astype(np.int64), pool_side)
    grey = s.uniform(n) < 0.6
    ss = np.array([4.0, 16.0, 32.0])[s.integers(0, 3, n)]
    sr = np.array([0.05, 0.2])[s.integers(0, 2, n)]
    x0 = np.floor(s.uniform(n) * (pool_side - side + 1)).astype(np.int64)
    y0 = np.floor(s.uniform(n) * (pool_side - side + 1)).astype(np.int64)
    return {"arrival_ms": arrival, "side": side, "C": np.where(grey, 1, 3), "sigma_s": ss, "sigma_r": sr,
            "x0": x0, "y0": y0}
