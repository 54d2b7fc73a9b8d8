// Launchers of the non-pixel lattice kernels: bounds, clear, compaction,
// key packing, finalize, neighbour table, blur, export helpers.
#include "gt_launch.h"

namespace gt {

// ---------------------------------------------------------------------------
// Per-plane bounds (for the key packing plan and the fixed-point value
// scales) of two groups of planes (A: nA planes at stride csA, B: nB planes
// at stride csB; n pixels each): min and max of the finite values and a
// non-finite indicator per plane and workgroup in partial[blk][plane][3].
__global__ __launch_bounds__(kBlock) void k_minmax(const float *__restrict__ pa, int64_t csa, int na,
                                                   const float *__restrict__ pb, int64_t csb, int nb, int n,
                                                   float *__restrict__ partial) {
    __shared__ float smn[kBlock / kWave], smx[kBlock / kWave], snf[kBlock / kWave];
    const int np = na + nb;
    for (int c = 0; c < np; c++) {
        float mn = INFINITY, mx = -INFINITY, nf = 0.0f;
        const float *pl = c < na ? pa + (int64_t)c * csa : pb + (int64_t)(c - na) * csb;
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
            const float v = pl[i];
            if (isfinite(v)) {
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
            } else {
                nf = 1.0f;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off, kWave));
            mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
            nf = fmaxf(nf, __shfl_xor(nf, off, kWave));
        }
        if (__lane_id() == 0) {
            smn[threadIdx.x / kWave] = mn;
            smx[threadIdx.x / kWave] = mx;
            snf[threadIdx.x / kWave] = nf;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kBlock / kWave; w++) {
                mn = fminf(mn, smn[w]);
                mx = fmaxf(mx, smx[w]);
                nf = fmaxf(nf, snf[w]);
            }
            float *o = partial + ((size_t)blockIdx.x * np + c) * 3;
            o[0] = mn;
            o[1] = mx;
            o[2] = nf;
        }
        __syncthreads();
    }
}

// Reduces partial[nblk][nplanes][3] to out[nplanes][3]: one workgroup per
// plane, the workgroup's threads striding over the blocks.
__global__ __launch_bounds__(kBlock) void k_minmax_final(const float *__restrict__ partial, int nblk, int nplanes,
                                                         float *__restrict__ out) {
    __shared__ float smn[kBlock / kWave], smx[kBlock / kWave], snf[kBlock / kWave];
    const int c = blockIdx.x;
    float mn = INFINITY, mx = -INFINITY, nf = 0.0f;
    for (int b = threadIdx.x; b < nblk; b += kBlock) {
        const float *o = partial + ((size_t)b * nplanes + c) * 3;
        mn = fminf(mn, o[0]);
        mx = fmaxf(mx, o[1]);
        nf = fmaxf(nf, o[2]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, off, kWave));
        mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
        nf = fmaxf(nf, __shfl_xor(nf, off, kWave));
    }
    if (__lane_id() == 0) {
        smn[threadIdx.x / kWave] = mn;
        smx[threadIdx.x / kWave] = mx;
        snf[threadIdx.x / kWave] = nf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; w++) {
            mn = fminf(mn, smn[w]);
            mx = fmaxf(mx, smx[w]);
            nf = fmaxf(nf, snf[w]);
        }
        out[c * 3 + 0] = mn;
        out[c * 3 + 1] = mx;
        out[c * 3 + 2] = nf;
    }
}

// Bounds of a batch: per guide channel the min/max of ref * ic (the scaled
// colour position, ic per image) and per image channel the min/max of the
// value, each with a non-finite indicator, as partial[blk][plane][3] for
// k_minmax_final.  A workgroup's tiles rise, so the image of its next tile
// is found by stepping forward from the last one (the last image whose first
// tile is at or before it), eight tiles' loads are in flight at a time, and
// the tiles are taken as runs of pixels (tile_w is not used), so a wave reads
// consecutive values (64 images of 30.8 M pixels: 488 -> 257 us against one
// 2-D tile and one binary search per step).
__global__ __launch_bounds__(kBlock) void k_bounds_batch(BatchArgs ba, int ntiles, int tile_w, int Cr, int C,
                                                         float *__restrict__ partial) {
    __shared__ float red[3][kBlock / kWave];
    const int np = Cr + C;
    constexpr int U = 8;
    for (int c = 0; c < np; c++) {
        float mn = INFINITY, mx = -INFINITY, nf = 0.0f;
        int img = 0;   // tile 0 starts image 0
        for (int t0 = blockIdx.x; t0 < ntiles; t0 += U * gridDim.x) {
            float v[U];
            bool have[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int tile = t0 + u * (int)gridDim.x;
                have[u] = false;
                v[u] = 0.0f;
                if (tile < ntiles) {
                    while (img + 1 < ba.count && ba.img[img + 1].tile0 <= tile) img++;
                    const BatchImage &bi = ba.img[img];
                    const int n = bi.W * bi.H;
                    // the bounds take every pixel once in any order: the
                    // image's k-th tile covers pixels [256 k, 256 k + 256)
                    // (its tiles number at least n / 256), so a wave reads
                    // 64 consecutive values
                    const int p = (tile - bi.tile0) * kBlock + (int)threadIdx.x;
                    if (p < n) {
                        have[u] = true;
                        v[u] = c < Cr ? bi.ref[(int64_t)c * n + p] * bi.ic : bi.im[(int64_t)(c - Cr) * n + p];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (!have[u]) continue;
                if (isfinite(v[u])) {
                    mn = fminf(mn, v[u]);
                    mx = fmaxf(mx, v[u]);
                } else {
                    nf = 1.0f;
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off, kWave));
            mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
            nf = fmaxf(nf, __shfl_xor(nf, off, kWave));
        }
        if (__lane_id() == 0) {
            red[0][threadIdx.x / kWave] = mn;
            red[1][threadIdx.x / kWave] = mx;
            red[2][threadIdx.x / kWave] = nf;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kBlock / kWave; w++) {
                mn = fminf(mn, red[0][w]);
                mx = fmaxf(mx, red[1][w]);
                nf = fmaxf(nf, red[2][w]);
            }
            float *o = partial + ((size_t)blockIdx.x * np + c) * 3;
            o[0] = mn;
            o[1] = mx;
            o[2] = nf;
        }
        __syncthreads();
    }
}

hipError_t launch_bounds_batch(const BatchArgs &ba, int ntiles, int tile_w, int Cr, int C, float *partial, int nblk,
                               float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_bounds_batch, dim3(nblk), dim3(kBlock), 0, st, ba, ntiles, tile_w, Cr, C, partial);
    hipLaunchKernelGGL(k_minmax_final, dim3(Cr + C), dim3(kBlock), 0, st, partial, nblk, Cr + C, out);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Value bounds of this call, for the fixed-point scales.  Per workgroup and
// value plane: the largest finite magnitude and a non-finite indicator, as
// partial[blk][plane][3] = {-max, max, nonfinite} (the layout k_minmax and
// k_bounds_batch write: min, max, indicator).  Four loads in flight per lane,
// 16-byte ones when the planes allow.
__global__ __launch_bounds__(kBlock) void k_value_bounds(const float *__restrict__ val, int64_t cs, int nval,
                                                         int n, float *__restrict__ partial) {
    __shared__ float smx[kBlock / kWave], snf[kBlock / kWave];
    const bool vec = ((uintptr_t)val % 16) == 0 && (cs % 4) == 0;
    const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * kBlock;
    for (int c = 0; c < nval; c++) {
        const float *pl = val + (int64_t)c * cs;
        float mx = 0.0f, nf = 0.0f;
        auto take = [&](float v) {
            const float a = fabsf(v);
            const bool fin = a <= 3.402823466e38f;   // false for NaN and Inf
            mx = fin ? fmaxf(mx, a) : mx;
            nf = fin ? nf : 1.0f;
        };
        int64_t done = 0;
        if (vec) {
            const float4 *p4 = reinterpret_cast<const float4 *>(pl);
            const int64_t n4 = n / 4;
            int64_t i = tid;
            for (; i + 3 * nth < n4; i += 4 * nth) {
                float4 q[4];
#pragma unroll
                for (int k = 0; k < 4; k++) q[k] = p4[i + k * nth];
#pragma unroll
                for (int k = 0; k < 4; k++) { take(q[k].x); take(q[k].y); take(q[k].z); take(q[k].w); }
            }
            for (; i < n4; i += nth) {
                const float4 q = p4[i];
                take(q.x); take(q.y); take(q.z); take(q.w);
            }
            done = n4 * 4;
        }
        for (int64_t i = done + tid; i < n; i += nth) take(pl[i]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
            nf = fmaxf(nf, __shfl_xor(nf, off, kWave));
        }
        if (__lane_id() == 0) {
            smx[threadIdx.x / kWave] = mx;
            snf[threadIdx.x / kWave] = nf;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kBlock / kWave; w++) {
                mx = fmaxf(mx, smx[w]);
                nf = fmaxf(nf, snf[w]);
            }
            float *o = partial + ((size_t)blockIdx.x * nval + c) * 3;
            o[0] = -mx;
            o[1] = mx;
            o[2] = nf;
        }
        __syncthreads();
    }
}

// This call's fixed-point scales from the per-workgroup bounds of the nval
// value planes (a workgroup that saw no finite value holds min > max); a
// homogeneous channel (homog) has magnitude 1.  For n contributions of
// magnitude <= vmax the scale is 2^p with p = min(62 - ceil(log2(n vmax
// 1.001)), 50 - ceil(log2(vmax 1.001))), clamped to [-120, 120]: every sum
// stays below 2^62 and every product below 2^50 (fixed_word's range).
// ceil(log2) is exact through frexp.  One workgroup, a wave per channel.
__device__ void value_scale_wg(const float *__restrict__ partial, int nblk, int nval, int homog, int n, int vd,
                               ValScale *__restrict__ out, Counters *ctr) {
    __shared__ unsigned s_nf;
    if (threadIdx.x == 0) s_nf = 0;
    // a narrow storage's per-pass magnitudes start from zero in every call
    for (int k = threadIdx.x; k < (kMaxDH + 2) * kMaxDH; k += kBlock) (&out->vmax_bits[0][0])[k] = 0u;
    __syncthreads();
    const int lane = __lane_id();
    for (int c = threadIdx.x / kWave; c < kMaxDH; c += kBlock / kWave) {
        float vmax = 0.0f, nf = 0.0f;
        if (c < nval) {
            // sixteen workgroups' bounds per lane in flight at once: 1024
            // partials are one round of loads, not sixteen
            constexpr int U = 16;
            for (int b0 = lane; b0 < nblk; b0 += U * kWave) {
                float lo[U], hi[U], fl[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int b = b0 + u * kWave;
                    if (b < nblk) {
                        const float *o = partial + ((size_t)b * nval + c) * 3;
                        lo[u] = o[0];
                        hi[u] = o[1];
                        fl[u] = o[2];
                    } else {
                        lo[u] = 1.0f;   // min > max: no finite value
                        hi[u] = 0.0f;
                        fl[u] = 0.0f;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (lo[u] <= hi[u]) vmax = fmaxf(vmax, fmaxf(-lo[u], hi[u]));
                    nf = fmaxf(nf, fl[u]);
                }
            }
        } else if (c == nval && homog) {
            vmax = 1.0f;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            vmax = fmaxf(vmax, __shfl_xor(vmax, off, kWave));
            nf = fmaxf(nf, __shfl_xor(nf, off, kWave));
        }
        if (lane == 0) {
            float s = 0.0f, inv = 0.0f;
            if (c < vd) {
                // every sum below 2^62 (n contributions of weight <= 1), every
                // single product below 2^50 (the splat's FP64 rounding of a
                // product to an integer is exact below 2^51)
                const double bound = fmax(1e-30, (double)n * (double)vmax * 1.001);
                int e = 0;
                double f = frexp(bound, &e);   // bound = f 2^e, f in [0.5, 1)
                const int lg = (f == 0.5) ? e - 1 : e;   // ceil(log2(bound))
                f = frexp(fmax(1e-30, (double)vmax * 1.001), &e);
                const int lg1 = (f == 0.5) ? e - 1 : e;
                const int p = min(120, max(-120, min(62 - lg, 50 - lg1)));
                s = ldexpf(1.0f, p);
                inv = ldexpf(1.0f, -p);
            }
            out->s[c] = s;
            out->inv_s[c] = inv;
            if (nf != 0.0f) atomicOr(&s_nf, 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out->nonfinite = s_nf;
        if (s_nf) atomicOr(&ctr->flags, kFlagNonFinite);
    }
}

// The table cleared (grid-stride), the counters zeroed, and this call's
// fixed-point scales computed by workgroup 0 from the value bounds' partials
// (the kernel before).  One launch for what needs no grid-wide order.
__global__ __launch_bounds__(kBlock) void k_clear_scale(ClearArgs cl, Counters *ctr, const float *__restrict__ partial,
                                                        int nblk, int nval, int homog, int n, int vd,
                                                        ValScale *__restrict__ out) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    const size_t t0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
    for (size_t i = t0; i < cl.nwords; i += stride) cl.tkeys[i] = kEmpty;
    for (size_t i = t0; i < cl.nacc; i += stride) cl.tacc[i] = 0ull;
    for (size_t i = t0; i < cl.nmarks; i += stride) cl.tnf[i] = 0u;
    for (size_t i = t0; i < cl.ntags / 4; i += stride) reinterpret_cast<uint32_t *>(cl.tags)[i] = 0u;
    if (blockIdx.x != 0) return;
    if (threadIdx.x == 0) {
        ctr->M = 0;
        ctr->flags = 0;
        ctr->filled = 0;
        ctr->probes = 0;
        ctr->inserts = 0;
        ctr->atomics = 0;
    }
    __syncthreads();
    value_scale_wg(partial, nblk, nval, homog, n, vd, out, ctr);
}

// In-place exclusive scan of n counts in one workgroup; the total goes to
// *total, or 0 when `flags` carries kFlagFull (the call is redone with a
// larger table, so everything after the compaction sees no vertices).
// A table holding more than `limit` vertices (the per-vertex buffers' size,
// below the table capacity only for a caller-bounded arena) counts as full.
__global__ __launch_bounds__(1024) void k_scan_1wg(uint32_t *__restrict__ a, int n, unsigned *total,
                                                   unsigned *flags, unsigned *filled, unsigned limit) {
    __shared__ uint32_t part[1024];
    const int chunk = (n + 1023) / 1024;
    const int lo = threadIdx.x * chunk;
    const int hi = min(n, lo + chunk);
    uint32_t s = 0;
    for (int i = lo; i < hi; i += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = i + q < hi ? a[i + q] : 0u;
#pragma unroll
        for (int q = 0; q < 8; q++) s += v[q];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (int i = lo; i < hi; i += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = i + q < hi ? a[i + q] : 0u;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (i + q < hi) a[i + q] = run;
            run += v[q];
        }
    }
    if (threadIdx.x == 1023 && total) {
        bool full = flags && (*flags & kFlagFull);
        if (flags && part[1023] > limit && !full) {
            atomicOr(flags, kFlagFull);
            full = true;
        }
        *total = full ? 0u : part[1023];
        if (filled) *filled = part[1023];
    }
}

// Strided view -> strided view copy, used to gather the non-dense views of
// JointBilateral's zero-sigma recursion into dense scratch images and back.
__global__ __launch_bounds__(kBlock) void k_view_copy(const float *__restrict__ src, int64_t sys, int64_t sts,
                                                      int64_t scs, float *__restrict__ dst, int64_t dys, int64_t dts,
                                                      int64_t dcs, int W, int H, int F, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int x = (int)(i % W);
    int64_t q = i / W;
    const int y = (int)(q % H);
    q /= H;
    const int t = (int)(q % F);
    const int64_t c = q / F;
    dst[c * dcs + t * dts + y * dys + x] = src[c * scs + t * sts + y * sys + x];
}


hipError_t launch_minmax(const float *pa, int64_t csa, int na, const float *pb, int64_t csb, int nb, int n,
                         float *partial, int nblk, float *out, hipStream_t st) {
    if (na + nb <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_minmax, dim3(nblk), dim3(kBlock), 0, st, pa, csa, na, pb, csb, nb, n, partial);
    hipLaunchKernelGGL(k_minmax_final, dim3(na + nb), dim3(kBlock), 0, st, partial, nblk, na + nb, out);
    return hipGetLastError();
}

static unsigned clear_blocks(const ClearArgs &cl) {
    size_t m = cl.nwords > cl.nacc ? cl.nwords : cl.nacc;
    if (cl.nmarks > m) m = cl.nmarks;
    unsigned g = blocks_for((int64_t)m, kBlock);
    if (g > 8192) g = 8192;
    return g == 0 ? 1u : g;
}

hipError_t launch_value_scale(const ClearArgs &cl, const float *val, int64_t cs, int nval, int homog, int n, int vd,
                              float *partial, int nblk, ValScale *out, Counters *ctr, hipStream_t st) {
    if (nval > 0 && n > 0)
        hipLaunchKernelGGL(k_value_bounds, dim3(nblk), dim3(kBlock), 0, st, val, cs, nval, n, partial);
    else
        nval = 0;
    hipLaunchKernelGGL(k_clear_scale, dim3(clear_blocks(cl)), dim3(kBlock), 0, st, cl, ctr, partial, nblk, nval,
                       homog, n, vd, out);
    return hipGetLastError();
}

hipError_t launch_value_scale_batch(const ClearArgs &cl, const BatchArgs &ba, int ntiles, int tile_w, int C, int n,
                                    int vd, float *partial, int nblk, ValScale *out, Counters *ctr, hipStream_t st) {
    if (C > 0 && ntiles > 0)
        hipLaunchKernelGGL(k_bounds_batch, dim3(nblk), dim3(kBlock), 0, st, ba, ntiles, tile_w, 0, C, partial);
    else
        C = 0;
    hipLaunchKernelGGL(k_clear_scale, dim3(clear_blocks(cl)), dim3(kBlock), 0, st, cl, ctr, partial, nblk, C, 1, n,
                       vd, out);
    return hipGetLastError();
}

hipError_t launch_compact(int words, const TableArgs &ta, uint32_t capacity, uint32_t *counts, uint32_t *slot_list,
                          uint64_t *sort_keys, bool packed, Counters *ctr, int limit, hipStream_t st) {
    const unsigned nblk = blocks_for(capacity, kCompactSlots);
    switch (words) {
    case 1: hipLaunchKernelGGL(k_count<1>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, nullptr); break;
    case 2: hipLaunchKernelGGL(k_count<2>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, nullptr); break;
    case 3: hipLaunchKernelGGL(k_count<3>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, nullptr); break;
    case 4: hipLaunchKernelGGL(k_count<4>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, nullptr); break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_scan_1wg, dim3(1), dim3(1024), 0, st, counts, (int)nblk, &ctr->M, &ctr->flags,
                       &ctr->filled, (unsigned)limit);
    switch (words) {
    case 1:
        if (packed) hipLaunchKernelGGL((k_write_list<1, true>), dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity,
                                       counts, slot_list, sort_keys);
        else hipLaunchKernelGGL((k_write_list<1, false>), dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity,
                                counts, slot_list, sort_keys);
        break;
    case 2:
        hipLaunchKernelGGL((k_write_list<2, false>), dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts,
                           slot_list, sort_keys);
        break;
    case 3:
        hipLaunchKernelGGL((k_write_list<3, false>), dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts,
                           slot_list, sort_keys);
        break;
    case 4:
        hipLaunchKernelGGL((k_write_list<4, false>), dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts,
                           slot_list, sort_keys);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_pack(int d, const TableArgs &ta, const uint32_t *slot_list, const Counters *ctr, int m_max,
                       const KeyPlan &kp, int pw, uint64_t *pk, hipStream_t st) {
    if (m_max <= 0) return hipSuccess;
#define GT_PACK(D) hipLaunchKernelGGL(k_pack<D>, dim3(blocks_for(m_max, kBlock)), dim3(kBlock), 0, st, ta, slot_list, \
                                      ctr, m_max, kp, pw, pk, (Counters *)ctr)
    GT_DISPATCH_D(d, GT_PACK)
#undef GT_PACK
    return hipGetLastError();
}

hipError_t launch_finalize(int words, int storage, const TableArgs &ta, const ValScale *vsc, const uint32_t *sorted,
                           const uint64_t *sorted_keys, const Counters *ctr, int m_max, int vd, int vs,
                           uint64_t *vkeys, void *vals, float *splat_copy, hipStream_t st) {
    if (m_max <= 0) return hipSuccess;
    const dim3 g(blocks_for(m_max, kBlock)), b(kBlock);
    // FP32 rows for every storage; a narrow one's first rounding is blur pass
    // 0, scaled by the largest magnitude tracked here
    unsigned *track = storage == 0 ? nullptr : const_cast<unsigned *>(&vsc->vmax_bits[0][0]);
    hipLaunchKernelGGL(k_finalize<float>, g, b, 0, st, ta, vsc, words, sorted, sorted_keys, ctr, vd, vs, vkeys,
                       (float *)vals, splat_copy, track);
    return hipGetLastError();
}

// Neighbour entries [0, M) of axis blockIdx.y set to {-1, -1} (the one-end
// neighbour form fills their second halves by scatter).
__global__ __launch_bounds__(kBlock) void k_nbr_fill(int2 *__restrict__ nbr, int stride, const Counters *ctr) {
    const unsigned M = ctr->M;
    u32x2 *row = reinterpret_cast<u32x2 *>(nbr) + (size_t)blockIdx.y * stride;
    const u32x2 none = {~0u, ~0u};
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < M; i += gridDim.x * kBlock)
        __builtin_nontemporal_store(none, row + i);
}

template <typename T>
static void launch_write_final_t(int words, const TableArgs &ta, uint32_t capacity, unsigned nblk,
                                 const uint32_t *offsets, const Counters *ctr, const ValScale *vsc, int vd, int vs,
                                 uint64_t *vkeys, void *vals, unsigned *track, const SelfScan &sc,
                                 const NbrInit &ninit, hipStream_t st) {
    T *v = (T *)vals;
#define GT_WF(W)                                                                                                 \
    hipLaunchKernelGGL((k_write_final<W, T>), dim3(nblk), dim3(kBlock), 0, st, ta, capacity, offsets, ctr, vsc,  \
                       vd, vs, vkeys, v, track, sc, ninit)
    switch (words) {
    case 1: GT_WF(1); break;
    case 2: GT_WF(2); break;
    case 3: GT_WF(3); break;
    default: GT_WF(4); break;
    }
#undef GT_WF
}

hipError_t launch_compact_final(int words, int storage, const TableArgs &ta, uint32_t capacity, uint32_t *counts,
                                uint32_t *offs, const ValScale *vsc, int vd, int vs, uint64_t *vkeys, void *vals,
                                Counters *ctr, int limit, const NbrInit &ninit, hipStream_t st) {
    if (words < 1 || words > 4) return hipErrorInvalidValue;
    if (ninit.nbr && (ninit.stride < limit || ninit.axes < 1 || ninit.axes > kMaxD + 1 || ninit.grid < 0))
        return hipErrorInvalidValue;
    const unsigned nblk = blocks_for(capacity, kCompactSlots);
    switch (words) {
    case 1: hipLaunchKernelGGL(k_count<1>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, ta.tags); break;
    case 2: hipLaunchKernelGGL(k_count<2>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, ta.tags); break;
    case 3: hipLaunchKernelGGL(k_count<3>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, ta.tags); break;
    default: hipLaunchKernelGGL(k_count<4>, dim3(nblk), dim3(kBlock), 0, st, ta.tkeys, capacity, counts, ta.tags); break;
    }
    SelfScan sc{nullptr, ctr, (int)nblk, (unsigned)limit};
    if (nblk <= (unsigned)kSelfScanChunks) sc.offs = offs;
    else
        hipLaunchKernelGGL(k_scan_1wg, dim3(1), dim3(1024), 0, st, counts, (int)nblk, &ctr->M, &ctr->flags,
                           &ctr->filled, (unsigned)limit);
    // FP32 rows for every storage (see launch_finalize)
    unsigned *track = storage == 0 ? nullptr : const_cast<unsigned *>(&vsc->vmax_bits[0][0]);
    launch_write_final_t<float>(words, ta, capacity, nblk, counts, ctr, vsc, vd, vs, vkeys, vals, track, sc, ninit,
                                st);
    if (ninit.nbr && ninit.grid > 0)
        hipLaunchKernelGGL(k_nbr_fill, dim3(blocks_for(ninit.grid, kBlock), ninit.axes), dim3(kBlock), 0, st,
                           ninit.nbr, ninit.stride, ctr);
    return hipGetLastError();
}

// Neighbour table by merging with the sorted keys (no hash lookups).
hipError_t launch_neighbors(int d, bool packed, const KeyPlan &kp, const uint64_t *vkeys, const Counters *ctr,
                            int m_max, int2 *nbr, int wpc, hipStream_t st) {
    if (m_max <= 0) return hipSuccess;
    const dim3 g(blocks_for(m_max, kNbrBlock), 2 * (d + 1)), b(kBlock);
    const unsigned pad = occupancy_pad(kNbrWindowWords * 8 + 256, wpc);
#define GT_NBR(D)                                                                                                \
    {                                                                                                            \
        if (packed) hipLaunchKernelGGL((k_nbr_merge<D, true>), g, b, pad, st, vkeys, kp, ctr, m_max, (int *)nbr);  \
        else hipLaunchKernelGGL((k_nbr_merge<D, false>), g, b, pad, st, vkeys, kp, ctr, m_max, (int *)nbr);        \
    }
    GT_DISPATCH_D(d, GT_NBR)
#undef GT_NBR
    return hipGetLastError();
}

// Workgroups along x for m_max vertices; a multiple of 8 with a part map,
// so that workgroup x of every row of a 2-D grid sits on XCD x % 8.
static unsigned vertex_blocks(int m_max, const PartMap &pm) {
    const unsigned g = blocks_for(m_max, kBlock);
    return pm.cpp ? (g + 7u) & ~7u : g;
}

hipError_t launch_neighbors_hash(int d, bool packed, const KeyPlan &kp, const uint64_t *vkeys, const Counters *ctr,
                                 int m_max, int m_grid, const TableArgs &ta, const PartMap &pm, bool per_vertex,
                                 bool half, int2 *nbr, hipStream_t st) {
    if (m_max <= 0) return hipSuccess;
    if (half && !per_vertex) return hipErrorInvalidValue;
    if (m_grid <= 0 || m_grid > m_max) m_grid = m_max;
    if (per_vertex) {
        const dim3 g1(vertex_blocks(m_grid, pm)), b1(kBlock);
#define GT_NBRV_H(D, P)                                                                                             \
    {                                                                                                               \
        if (half) hipLaunchKernelGGL((k_nbr_hash_vertex<D, P, true>), g1, b1, 0, st, vkeys, kp, ctr, m_max, ta, pm, \
                                     (int *)nbr);                                                                   \
        else hipLaunchKernelGGL((k_nbr_hash_vertex<D, P, false>), g1, b1, 0, st, vkeys, kp, ctr, m_max, ta, pm,     \
                                (int *)nbr);                                                                        \
    }
#define GT_NBRV(D)                 \
    {                              \
        if (packed) GT_NBRV_H(D, true) \
        else GT_NBRV_H(D, false)   \
    } break;
        switch (d) {
        case 1: GT_NBRV(1) case 2: GT_NBRV(2) case 3: GT_NBRV(3) case 4: GT_NBRV(4) case 5: GT_NBRV(5)
        case 6: GT_NBRV(6) case 7: GT_NBRV(7) case 8: GT_NBRV(8) case 9: GT_NBRV(9) case 10: GT_NBRV(10)
        case 11: GT_NBRV(11) case 12: GT_NBRV(12) case 13: GT_NBRV(13) case 14: GT_NBRV(14) case 15: GT_NBRV(15)
        case 16: GT_NBRV(16)
        default: return hipErrorInvalidValue;
        }
#undef GT_NBRV
#undef GT_NBRV_H
        return hipGetLastError();
    }
    const dim3 g(vertex_blocks(m_grid, pm), d + 1), b(kBlock);
#define GT_NBRH(D)                                                                                                  \
    {                                                                                                               \
        if (packed)                                                                                                 \
            hipLaunchKernelGGL((k_nbr_hash<D, true>), g, b, 0, st, vkeys, kp, ctr, m_max, ta, pm, (int *)nbr);      \
        else hipLaunchKernelGGL((k_nbr_hash<D, false>), g, b, 0, st, vkeys, kp, ctr, m_max, ta, pm, (int *)nbr);    \
    }
    GT_DISPATCH_D(d, GT_NBRH)
#undef GT_NBRH
    return hipGetLastError();
}

template <typename TI, typename TO, bool SCALED>
static hipError_t launch_blur_t(int vs, const int2 *nbr, const Counters *ctr, int m_max, const PartMap &pm,
                                const void *in, void *out, const unsigned *in_bits, const unsigned *out_bits,
                                unsigned *out_max, hipStream_t st) {
    const dim3 g(vertex_blocks(m_max, pm)), b(kBlock);
    const TI *i = (const TI *)in;
    TO *o = (TO *)out;
#define GT_BLUR(VS) hipLaunchKernelGGL((k_blur<TI, TO, VS, SCALED>), g, b, 0, st, nbr, ctr, pm, i, o, in_bits, \
                                       out_bits, out_max)
    switch (vs) {
    case 2: GT_BLUR(2); break;
    case 4: GT_BLUR(4); break;
    case 8: GT_BLUR(8); break;
    case 16: GT_BLUR(16); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_BLUR
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_blur_narrow(int vs, int pass, const ValScale *vsc, const int2 *nbr, const Counters *ctr,
                                     int m_max, const PartMap &pm, const void *in, void *out, hipStream_t st) {
    const unsigned *out_bits = &vsc->vmax_bits[pass][0];
    unsigned *out_max = const_cast<unsigned *>(&vsc->vmax_bits[pass + 1][0]);
    if (pass == 0)   // the finalize's FP32 rows, unscaled
        return launch_blur_t<float, T, true>(vs, nbr, ctr, m_max, pm, in, out, nullptr, out_bits, out_max, st);
    return launch_blur_t<T, T, true>(vs, nbr, ctr, m_max, pm, in, out, &vsc->vmax_bits[pass - 1][0], out_bits,
                                     out_max, st);
}

hipError_t launch_blur2(int vs, const int2 *nbr_j, const int2 *nbr_k, const Counters *ctr, int m_max,
                        const PartMap &pm, const float *in, float *out, hipStream_t st) {
    if (m_max <= 0) return hipSuccess;
    const dim3 g(vertex_blocks(m_max, pm)), b(kBlock);
    if (vs == 2) hipLaunchKernelGGL(k_blur2<2>, g, b, 0, st, nbr_j, nbr_k, ctr, pm, in, out);
    else if (vs == 4) hipLaunchKernelGGL(k_blur2<4>, g, b, 0, st, nbr_j, nbr_k, ctr, pm, in, out);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_blur(int storage, int vs, int pass, const ValScale *vsc, const int2 *nbr, const Counters *ctr,
                       int m_max, const PartMap &pm, const void *in, void *out, hipStream_t st) {
    if (m_max <= 0) return hipSuccess;
    if (pass < 0 || pass > kMaxD) return hipErrorInvalidValue;
    if (storage == 0)
        return launch_blur_t<float, float, false>(vs, nbr, ctr, m_max, pm, in, out, nullptr, nullptr, nullptr, st);
    if (storage == 1) return launch_blur_narrow<__hip_bfloat16>(vs, pass, vsc, nbr, ctr, m_max, pm, in, out, st);
    return launch_blur_narrow<__half>(vs, pass, vsc, nbr, ctr, m_max, pm, in, out, st);
}

hipError_t launch_unpack_keys(int d, bool packed, const KeyPlan &kp, const uint64_t *vkeys, int M, int16_t *keys,
                              hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const dim3 g(blocks_for(M, kBlock)), b(kBlock);
#define GT_UNPACK(D)                                                                                       \
    {                                                                                                      \
        if (packed) hipLaunchKernelGGL((k_unpack_keys<D, true>), g, b, 0, st, kp, vkeys, M, keys);         \
        else hipLaunchKernelGGL((k_unpack_keys<D, false>), g, b, 0, st, kp, vkeys, M, keys);               \
    }
    GT_DISPATCH_D(d, GT_UNPACK)
#undef GT_UNPACK
    return hipGetLastError();
}

// One thread per image: lower bound of tag i in the sorted one-word keys.
__global__ __launch_bounds__(kBlock) void k_tag_starts(const uint64_t *__restrict__ vkeys, int M, int coord_bits,
                                                       int count, uint32_t *__restrict__ starts) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i > count) return;
    int lo = 0, hi = M;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if ((int64_t)(vkeys[mid] >> coord_bits) < (int64_t)i) lo = mid + 1;
        else hi = mid;
    }
    starts[i] = (uint32_t)lo;
}

hipError_t launch_tag_starts(const uint64_t *vkeys, int M, int coord_bits, int count, uint32_t *starts,
                             hipStream_t st) {
    if (count < 0 || M < 0 || coord_bits < 1 || coord_bits > 63) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tag_starts, dim3(blocks_for(count + 1, kBlock)), dim3(kBlock), 0, st, vkeys, M, coord_bits,
                       count, starts);
    return hipGetLastError();
}

hipError_t launch_store_to_f32(int storage, const void *vals, const unsigned *sbits, int M, int vd, int vs,
                               float *out, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const dim3 g(blocks_for(M, kBlock)), b(kBlock);
    if (storage == 0)
        hipLaunchKernelGGL(k_store_to_f32<float>, g, b, 0, st, (const float *)vals, sbits, M, vd, vs, out);
    else if (storage == 1)
        hipLaunchKernelGGL(k_store_to_f32<__hip_bfloat16>, g, b, 0, st, (const __hip_bfloat16 *)vals, sbits, M, vd,
                           vs, out);
    else hipLaunchKernelGGL(k_store_to_f32<__half>, g, b, 0, st, (const __half *)vals, sbits, M, vd, vs, out);
    return hipGetLastError();
}

hipError_t launch_view_copy(const float *src, int64_t sys, int64_t sts, int64_t scs, float *dst, int64_t dys,
                            int64_t dts, int64_t dcs, int W, int H, int F, int C, hipStream_t st) {
    const int64_t total = (int64_t)W * H * F * C;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_view_copy, dim3(blocks_for(total, kBlock)), dim3(kBlock), 0, st, src, sys, sts, scs, dst, dys,
                       dts, dcs, W, H, F, total);
    return hipGetLastError();
}

} // namespace gt
