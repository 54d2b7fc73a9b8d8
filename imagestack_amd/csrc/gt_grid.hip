// Bilateral grid (the GRID method of GaussTransform::apply,
// src/GaussTransform.cpp:227-297, on src/DenseGrid.h with 5 taps), for
// gfx950.  Positions scaled by sf = sqrtf(1/3 + 1) index a dense grid of
// ceil(max - min) + 1 cells per dimension with vd interleaved values.
//
//   k_grid_splat  : one thread per pixel: cell indices and fractions in FP32
//                   exactly as the reference, 2^D corner weights multiplied in
//                   dimension order, and exact 64-bit fixed-point atomics (the
//                   sums do not depend on the order of the atomics).
//   k_grid_finish : fixed point -> FP32.
//   k_grid_blur   : one element per thread, one of the 2 rounds of the
//                   [1/4 1/2 1/4] blur along one dimension; each output is the
//                   reference's in-place row recurrence written out for that
//                   element (it only ever reads the row's previous values),
//                   with the same FP32 roundings and the same double-precision
//                   step for the last element of a row.
//   k_grid_slice  : one thread per pixel: the same corners in the same order,
//                   FP32 sums, then (JointBilateral) division by the
//                   homogeneous channel.
// A corner past the end of a dimension has weight exactly 0 (the fraction is
// 0 there) and is skipped.
#include "gt_launch.h"

#include <algorithm>
#include <utility>

namespace gt {

// Cell and fractions of pixel p (src/DenseGrid.h:300-305).  The bounds are
// the minimum and maximum of these same products, so every index lies in
// its dimension; a pixel that does not (which would take a bound computed
// from other values) is reported as outside and skipped.
template <int D, bool BILAT>
__device__ __forceinline__ bool grid_cell(const PixSrc &src, int p, const GridArgs &g, int64_t &base,
                                          float (&F)[D], float (&Fi)[D], int (&I)[D]) {
    float pos[D];
    load_position<D, BILAT>(src, p, pos);
    base = 0;
    bool in = true;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const float f = pos[i] * g.sf - g.mn[i];
        I[i] = (short)floorf(f);
        F[i] = f - (float)I[i];
        Fi[i] = 1 - F[i];
        in = in && I[i] >= 0 && I[i] < g.size[i];
        base += (int64_t)I[i] * g.stride[i];
    }
    return in;
}

// Splat, one pixel tile per workgroup (tile_w x 256/tile_w pixels of an
// image, or 256 consecutive pixels).  The tile's cells lie in the box
// spanned by its pixels' cell indices (+1 for the upper corners); when that
// box holds at most kGridLds values the tile pre-reduces its contributions in
// LDS with 64-bit integer atomics and adds each non-zero box value to the
// grid once; otherwise every contribution goes to the grid directly.  Both
// are exact fixed-point sums, so the result is the same either way.
constexpr int kGridLds = 4096;

template <int D, bool BILAT>
__global__ __launch_bounds__(kBlock) void k_grid_splat(PixSrc src, int n, int row_w, int tile_w, GridArgs g,
                                                       ValScale vsc, unsigned long long *__restrict__ acc) {
    __shared__ unsigned long long lacc[kGridLds];
    __shared__ int blo[D], bhi[D];
    const int tid = threadIdx.x;
    if (tid < D) {
        blo[tid] = 0x7fffffff;
        bhi[tid] = -1;
    }
    const int p = pixel_index(blockIdx.x, tid, n, row_w, tile_w);
    int64_t base = 0;
    float F[D], Fi[D];
    int I[D];
    const bool in = p >= 0 && grid_cell<D, BILAT>(src, p, g, base, F, Fi, I);
    float v[kMaxVD], vorig[kMaxVD];
#pragma unroll
    for (int k = 0; k < kMaxVD; k++) v[k] = 0.0f;
    if (in) load_values<kMaxVD>(src, p, v);
    // NaN/Inf values (the reference adds w * v in FP32, src/DenseGrid.h):
    // the fixed-point sums take them as 0 and the corner cells get a mark of
    // the product's class instead, which k_grid_finish turns into NaN / +-Inf
    unsigned nfmask = 0;
    if (g.marks) {
#pragma unroll
        for (int k = 0; k < kMaxVD; k++) vorig[k] = v[k];
        nfmask = split_nonfinite<kMaxVD>(vorig, g.vd, v);
    }
    __syncthreads();
    if (in) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            atomicMin(&blo[i], I[i]);
            atomicMax(&bhi[i], min(I[i] + 1, g.size[i] - 1));
        }
    }
    __syncthreads();
    // the box: extents and value strides (lstride[0] = vd)
    int lo[D], lstride[D];
    int64_t vol = g.vd;
    bool any = true;
#pragma unroll
    for (int i = 0; i < D; i++) {
        lo[i] = blo[i];
        const int e = bhi[i] - blo[i] + 1;
        any = any && e > 0;
        lstride[i] = (int)vol;
        vol *= e > 0 ? e : 1;
    }
    if (!any) return;   // no pixel of the tile lies in the grid (uniform)
    const bool local = vol <= kGridLds;   // uniform over the workgroup
    if (local) {
        for (int e = tid; e < (int)vol; e += kBlock) lacc[e] = 0ull;
        __syncthreads();
    }
    if (in) {
        // Lane i starts at corner i mod 2^D: neighbouring pixels share cells,
        // and visiting different corners in the same instruction spreads the
        // LDS atomics over more addresses (the sums are exact, so the order
        // does not matter).
        const int rot = tid & ((1 << D) - 1);
        for (int cc = 0; cc < (1 << D); cc++) {
            const int c = (cc + rot) & ((1 << D) - 1);
            float w = 1;
            int64_t idx = base;
            int li = 0;
            bool inside = true;
#pragma unroll
            for (int j = 0; j < D; j++) {
                const int bit = (c >> j) & 1;
                if (bit) {
                    idx += g.stride[j];
                    w *= F[j];
                    inside = inside && I[j] + 1 < g.size[j];
                } else {
                    w *= Fi[j];
                }
                li += (I[j] + bit - lo[j]) * lstride[j];
            }
            if (!inside) continue;
            if (nfmask) {
                for (int k = 0; k < g.vd; k++) {
                    if (!((nfmask >> k) & 1u)) continue;
                    const float prod = w * vorig[k];
                    atomicOr(g.marks + idx + k, (prod != prod) ? 3u : (prod > 0.0f ? 1u : 2u));
                }
            }
#pragma unroll
            for (int k = 0; k < kMaxVD; k++) {
                if (k >= g.vd) continue;
                const unsigned long long q = to_fixed(w * v[k], vsc.s[k]);
                if (local) atomicAdd(&lacc[li + k], q);
                else atomicAdd(acc + idx + k, q);
            }
        }
    }
    if (!local) return;
    __syncthreads();
    // flush: box value e -> grid element
    for (int e = tid; e < (int)vol; e += kBlock) {
        const unsigned long long q = lacc[e];
        if (q == 0ull) continue;
        int r = e;
        const int k = r % g.vd;
        r /= g.vd;
        int64_t idx = k;
#pragma unroll
        for (int j = 0; j < D; j++) {
            const int ext = bhi[j] - blo[j] + 1;
            idx += (int64_t)(lo[j] + r % ext) * g.stride[j];
            r /= ext;
        }
        atomicAdd(acc + idx, q);
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_zero(unsigned long long *__restrict__ acc, int64_t total,
                                                      uint32_t *__restrict__ marks) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
        acc[i] = 0ull;
        if (marks) marks[i] = 0u;
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_finish(const unsigned long long *__restrict__ acc, int64_t total,
                                                        int vd, ValScale vsc, const uint32_t *__restrict__ marks,
                                                        float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    const int k = (int)(e % vd);
    float o = (float)((double)(long long)acc[e] * (double)vsc.inv_s[k]);
    if (marks) {
        const uint32_t m = marks[e];
        o = m == 3u ? __builtin_nanf("") : m == 1u ? INFINITY : m == 2u ? -INFINITY : o;
    }
    out[e] = o;
}

// One round along a dimension with element stride s and n cells:
//   first   : t1 = x/2
//   others  : t1 = 0.5 * (prev + x)          (the previous step's average)
//   not last: out = 0.5 * (t1 + 0.5 * (x + next))
//   last    : out = 0.5 * (t1 + 0.5 * x), in double (src/DenseGrid.h:268-271)
__global__ __launch_bounds__(kBlock) void k_grid_blur(const float *__restrict__ in, float *__restrict__ out,
                                                      int64_t total, int64_t s, int n) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    const int i = (int)((e / s) % n);
    const float x = in[e];
    const float t1 = i == 0 ? x / 2 : 0.5f * (in[e - s] + x);
    float r;
    if (i < n - 1) {
        const float t2 = 0.5f * (x + in[e + s]);
        r = 0.5f * (t1 + t2);
    } else {
        r = (float)(0.5 * ((double)t1 + 0.5 * (double)x));
    }
    out[e] = r;
}

template <int D, bool BILAT>
__global__ __launch_bounds__(kBlock) void k_grid_slice(PixSrc src, int n, GridArgs g,
                                                       const float *__restrict__ grid, OutDst dst) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    int64_t base;
    float F[D], Fi[D];
    int I[D];
    const bool in = grid_cell<D, BILAT>(src, p, g, base, F, Fi, I);
    float acc[kMaxVD];
#pragma unroll
    for (int k = 0; k < kMaxVD; k++) acc[k] = 0.0f;
    if (!in) {
        write_pixel<kMaxVD>(dst, p, g.vd, acc);
        return;
    }
    for (int c = 0; c < (1 << D); c++) {
        float w = 1;
        int64_t idx = base;
        bool inside = true;
#pragma unroll
        for (int j = 0; j < D; j++) {
            if (c & (1 << j)) {
                idx += g.stride[j];
                w *= F[j];
                inside = inside && I[j] + 1 < g.size[j];
            } else {
                w *= Fi[j];
            }
        }
        if (!inside) continue;
#pragma unroll
        for (int k = 0; k < kMaxVD; k++)
            if (k < g.vd) acc[k] = acc[k] + w * grid[idx + k];
    }
    write_pixel<kMaxVD>(dst, p, g.vd, acc);
}

#define GT_DISPATCH_GRID(d, MACRO)                                                \
    switch (d) {                                                                  \
    case 1: MACRO(1); break; case 2: MACRO(2); break; case 3: MACRO(3); break;    \
    case 4: MACRO(4); break; case 5: MACRO(5); break; case 6: MACRO(6); break;    \
    default: return hipErrorInvalidValue;                                         \
    }

hipError_t launch_grid(int d, bool bilat, const PixSrc &splat, int n, int row_w, const PixSrc &slice, int n_slice,
                       const OutDst &dst, const GridArgs &g, int64_t cells, const ValScale &vsc,
                       unsigned long long *acc, float *a, float *b, hipEvent_t *ev, hipStream_t st) {
    const int64_t total = cells * g.vd;
    const unsigned zb = (unsigned)std::min<int64_t>(8192, (total + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_grid_zero, dim3(zb > 0 ? zb : 1), dim3(kBlock), 0, st, acc, total, g.marks);
    if (ev) (void)hipEventRecord(ev[0], st);
    int tile_w = row_w > 0 ? 16 : 0;   // 16 x 16 pixel tiles of an image, else 256 consecutive pixels
    const int64_t ntiles = pixel_blocks(n, row_w, tile_w);
    const dim3 gp((unsigned)ntiles), gs((n_slice + kBlock - 1) / kBlock), blk(kBlock);
#define GT_GSPLAT(D)                                                                                          \
    {                                                                                                         \
        if (bilat) hipLaunchKernelGGL((k_grid_splat<D, true>), gp, blk, 0, st, splat, n, row_w, tile_w, g, vsc, \
                                      acc);                                                                   \
        else hipLaunchKernelGGL((k_grid_splat<D, false>), gp, blk, 0, st, splat, n, row_w, tile_w, g, vsc,    \
                                acc);                                                                         \
    }
    if (n > 0) GT_DISPATCH_GRID(d, GT_GSPLAT)
#undef GT_GSPLAT
    if (ev) (void)hipEventRecord(ev[1], st);
    const dim3 ge((unsigned)((total + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_grid_finish, ge, blk, 0, st, acc, total, g.vd, vsc, g.marks, a);
    if (ev) (void)hipEventRecord(ev[2], st);
    if (ev) (void)hipEventRecord(ev[3], st);
    float *cur = a, *nxt = b;
    for (int j = 0; j < d; j++)
        for (int it = 0; it < 2; it++) {
            hipLaunchKernelGGL(k_grid_blur, ge, blk, 0, st, cur, nxt, total, g.stride[j], g.size[j]);
            std::swap(cur, nxt);
        }
    if (ev) (void)hipEventRecord(ev[4], st);
#define GT_GSLICE(D)                                                                                        \
    {                                                                                                       \
        if (bilat) hipLaunchKernelGGL((k_grid_slice<D, true>), gs, blk, 0, st, slice, n_slice, g, cur, dst); \
        else hipLaunchKernelGGL((k_grid_slice<D, false>), gs, blk, 0, st, slice, n_slice, g, cur, dst);     \
    }
    if (n_slice > 0) GT_DISPATCH_GRID(d, GT_GSLICE)
#undef GT_GSLICE
    if (ev) (void)hipEventRecord(ev[5], st);
    return hipGetLastError();
}

} // namespace gt
