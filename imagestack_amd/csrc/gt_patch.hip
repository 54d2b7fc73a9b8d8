// Patch projection for non-local means on gfx950: the guide image of
// NLMeans::apply / NLMeans3D::apply (src/GaussTransform.cpp:867-874, 934-941),
// Convolve::apply(im, filters, Zero, Multiply::Inner) (src/Convolve.cpp:
// 248-270) with a bank of K*C filter planes from PatchPCA (src/Statistics.cpp:
// 1710-1770).
//
// Output channel k at pixel p is  sum_{c=0..C-1} v_kc  (added in c order to
// 0) where v_kc is the zero-boundary convolution of image channel c with
// filter plane k*C + c (convolveSingle, :151-173): taps visited in t, y, x
// order, the filter read at (xoff-dx, yoff-dy, toff-dt), taps outside the
// image skipped, every product and sum rounded to FP32 separately (the
// library is built with -ffp-contract=off).  This is the reference's order
// exactly, so the guide is bit-identical to the reference's.
//
//   k_patch_weights : filters [k][c][tap] -> [c][tap][k], so that the K
//                     weights of one tap are contiguous (one wide scalar load).
//   k_patch_project : one 16x16 pixel tile per workgroup (one pixel per lane),
//                     one (channel, frame) slab of the tile plus its halo
//                     staged in LDS at a time; weights are wave-uniform and
//                     come through the scalar cache.  Tiles whose halo lies
//                     inside the image skip the per-tap bounds tests.
#include "gt_internal.h"

namespace gt {
namespace {

constexpr int kPatchTile = 16;
constexpr int kPatchThreads = kPatchTile * kPatchTile;

__global__ __launch_bounds__(256) void k_patch_weights(const float *__restrict__ filt, int C, int K, int fsz,
                                                      float *__restrict__ wt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)K * C * fsz;
    if (i >= total) return;
    // destination order: c, tap, k
    const int k = (int)(i % K);
    const int64_t r = i / K;
    const int tap = (int)(r % fsz);
    const int c = (int)(r / fsz);
    wt[i] = filt[((int64_t)k * C + c) * fsz + tap];
}

template <int K>
__global__ __launch_bounds__(kPatchThreads) void k_patch_project(const float *__restrict__ im, int W, int H, int F,
                                                                 int C, const float *__restrict__ wt, int fw, int fh,
                                                                 int ft, float *__restrict__ out) {
    extern __shared__ float slab[];   // (16 + fw - 1) x (16 + fh - 1)
    const int sw = kPatchTile + fw - 1, sh = kPatchTile + fh - 1;
    const int xoff = (fw - 1) / 2, yoff = (fh - 1) / 2, toff = (ft - 1) / 2;
    const int x0 = blockIdx.x * kPatchTile, y0 = blockIdx.y * kPatchTile, t = blockIdx.z;
    const int tx = threadIdx.x % kPatchTile, ty = threadIdx.x / kPatchTile;
    const int x = x0 + tx, y = y0 + ty;
    const int64_t plane = (int64_t)W * H, n = plane * F;
    const int fsz = fw * fh * ft;
    const bool interior = x0 - xoff >= 0 && y0 - yoff >= 0 && x0 + kPatchTile - 1 + xoff < W &&
                          y0 + kPatchTile - 1 + yoff < H;
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.0f;
    for (int c = 0; c < C; c++) {
        float v[K];
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = 0.0f;
        for (int dt = -toff; dt <= toff; dt++) {
            const int tt = t + dt;
            if (tt < 0 || tt >= F) continue;   // uniform over the workgroup
            const float *src = im + (int64_t)c * n + (int64_t)tt * plane;
            __syncthreads();
            for (int i = threadIdx.x; i < sw * sh; i += kPatchThreads) {
                const int sx = x0 - xoff + i % sw, sy = y0 - yoff + i / sw;
                slab[i] = (sx >= 0 && sx < W && sy >= 0 && sy < H) ? src[(int64_t)sy * W + sx] : 0.0f;
            }
            __syncthreads();
            // weights of this (channel, frame offset): tap index (xoff-dx) + fw*(yoff-dy)
            const float *wc = wt + ((int64_t)c * fsz + (int64_t)fw * fh * (toff - dt)) * K;
            if (interior) {
                for (int dy = -yoff; dy <= yoff; dy++) {
                    const float *row = slab + (ty + yoff + dy) * sw + tx + xoff;
                    const float *wr = wc + (int64_t)fw * (yoff - dy) * K;
                    for (int dx = -xoff; dx <= xoff; dx++) {
                        const float a = row[dx];
                        const float *w = wr + (xoff - dx) * K;
#pragma unroll
                        for (int k = 0; k < K; k++) v[k] = v[k] + a * w[k];
                    }
                }
            } else {
                for (int dy = -yoff; dy <= yoff; dy++) {
                    const bool yin = y + dy >= 0 && y + dy < H;
                    const float *row = slab + (ty + yoff + dy) * sw + tx + xoff;
                    const float *wr = wc + (int64_t)fw * (yoff - dy) * K;
                    for (int dx = -xoff; dx <= xoff; dx++) {
                        const float a = row[dx];
                        const float *w = wr + (xoff - dx) * K;
                        if (yin && x + dx >= 0 && x + dx < W) {
#pragma unroll
                            for (int k = 0; k < K; k++) v[k] = v[k] + a * w[k];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = acc[k] + v[k];
    }
    if (x < W && y < H) {
        float *o = out + (int64_t)t * plane + (int64_t)y * W + x;
#pragma unroll
        for (int k = 0; k < K; k++) o[(int64_t)k * n] = acc[k];
    }
}

template <int K>
hipError_t launch_k(const float *im, int W, int H, int F, int C, const float *wt, int fw, int fh, int ft, float *out,
                    hipStream_t s) {
    const dim3 grid((W + kPatchTile - 1) / kPatchTile, (H + kPatchTile - 1) / kPatchTile, F);
    const size_t lds = sizeof(float) * (kPatchTile + fw - 1) * (kPatchTile + fh - 1);
    hipLaunchKernelGGL(k_patch_project<K>, grid, dim3(kPatchThreads), lds, s, im, W, H, F, C, wt, fw, fh, ft, out);
    return hipGetLastError();
}

} // namespace

size_t patch_lds_bytes(int fw, int fh) { return sizeof(float) * (kPatchTile + fw - 1) * (kPatchTile + fh - 1); }

hipError_t launch_patch_project(const float *im, int W, int H, int F, int C, const float *filters, int fw, int fh,
                                int ft, int K, float *wt, float *out, hipStream_t s) {
    const int fsz = fw * fh * ft;
    const int64_t total = (int64_t)K * C * fsz;
    hipLaunchKernelGGL(k_patch_weights, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, filters, C, K, fsz,
                       wt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (K) {
    case 1: return launch_k<1>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 2: return launch_k<2>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 3: return launch_k<3>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 4: return launch_k<4>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 5: return launch_k<5>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 6: return launch_k<6>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 7: return launch_k<7>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 8: return launch_k<8>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 9: return launch_k<9>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 10: return launch_k<10>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 11: return launch_k<11>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 12: return launch_k<12>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 13: return launch_k<13>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 14: return launch_k<14>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 15: return launch_k<15>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    case 16: return launch_k<16>(im, W, H, F, C, wt, fw, fh, ft, out, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace gt
