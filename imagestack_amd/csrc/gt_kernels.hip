// Lattice kernels for gfx950: splat (elevate + hash insert + accumulate),
// compaction, key packing for the radix sort, finalize (sorted ids), neighbour
// table, blur passes and slice.  Launchers at the bottom dispatch the runtime
// lattice dimension d to the compile-time template parameter D.
#include "gt_device.h"
#include "gt_internal.h"

namespace gt {

// ---------------------------------------------------------------------------
// value rows: vertex-major, `VS` values per vertex (VS a power of two >= vd)
template <typename T> __device__ __forceinline__ float to_f32(T v);
template <> __device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f32<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <> __device__ __forceinline__ float to_f32<__half>(__half v) { return __half2float(v); }
template <typename T> __device__ __forceinline__ T from_f32(float v);
template <> __device__ __forceinline__ float from_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ __hip_bfloat16 from_f32<__hip_bfloat16>(float v) { return __float2bfloat16(v); }
template <> __device__ __forceinline__ __half from_f32<__half>(float v) { return __float2half(v); }

template <typename T, int VS>
struct alignas((sizeof(T) * VS) < 16 ? (sizeof(T) * VS) : 16) Row {
    T v[VS];
};

template <typename T, int VS>
__device__ __forceinline__ void load_row(const T *base, int64_t i, float (&o)[kMaxVD]) {
    const Row<T, VS> r = reinterpret_cast<const Row<T, VS> *>(base)[i];
#pragma unroll
    for (int c = 0; c < VS; c++) o[c] = to_f32<T>(r.v[c]);
}

template <typename T, int VS>
__device__ __forceinline__ void store_row(T *base, int64_t i, const float (&o)[kMaxVD]) {
    Row<T, VS> r;
#pragma unroll
    for (int c = 0; c < VS; c++) r.v[c] = from_f32<T>(o[c]);
    reinterpret_cast<Row<T, VS> *>(base)[i] = r;
}

template <typename T>
__device__ __forceinline__ void load_row_rt(const T *base, int64_t i, int vs, float (&o)[kMaxVD]) {
    switch (vs) {
    case 1: load_row<T, 1>(base, i, o); break;
    case 2: load_row<T, 2>(base, i, o); break;
    case 4: load_row<T, 4>(base, i, o); break;
    case 8: load_row<T, 8>(base, i, o); break;
    default: load_row<T, 16>(base, i, o); break;
    }
}

// ---------------------------------------------------------------------------
// pixel sources

// Pixel index of thread `tid` of workgroup `blk`: plain linear order, or a
// tile_w x (B/tile_w) tile of an image with rows of row_w pixels.
__device__ __forceinline__ int pixel_index(int blk, int tid, int B, int n, int row_w, int tile_w) {
    if (tile_w <= 0 || row_w <= 0) {
        const int64_t p = (int64_t)blk * B + tid;
        return p < n ? (int)p : -1;
    }
    const int tiles_x = (row_w + tile_w - 1) / tile_w;
    const int th = B / tile_w;
    const int bx = blk % tiles_x, by = blk / tiles_x;
    const int x = bx * tile_w + tid % tile_w;
    const int y = by * th + tid / tile_w;
    const int rows = n / row_w;
    return (x < row_w && y < rows) ? y * row_w + x : -1;
}

template <int D, bool BILAT>
__device__ __forceinline__ void load_position(const PixSrc &s, int p, float (&pos)[D]) {
    if constexpr (BILAT) {
        // (ref_c * (1.0f/sigma_r), x*(1.0f/sigma_x), y*(1.0f/sigma_y),
        //  t*(1.0f/sigma_t)) for the filtered axes, in that order
        // (src/GaussTransform.cpp:654-666).
        const int x = p % s.W;
        const int q = p / s.W;
        const int y = q % s.H;
        const int t = q / s.H;
        const float xs = (float)x * s.ix, ys = (float)y * s.iy, ts = (float)t * s.it;
        const float a0 = s.fx ? xs : (s.fy ? ys : ts);
        const float a1 = s.fx ? (s.fy ? ys : ts) : ts;
#pragma unroll
        for (int c = 0; c < D; c++) {
            float v;
            if (c < s.cr) {
                v = s.ref[(int64_t)c * s.ref_cs + p] * s.ic;
            } else {
                const int k = c - s.cr;
                v = (k == 0) ? a0 : (k == 1 ? a1 : ts);
            }
            pos[c] = v;
        }
    } else {
        // pos[c] = splat(x,y,t,c) * invSigma[c] (src/GaussTransform.cpp:176-178)
#pragma unroll
        for (int c = 0; c < D; c++) pos[c] = s.pos[(int64_t)c * s.pos_cs + p] * s.inv[c];
    }
}

__device__ __forceinline__ void load_values(const PixSrc &s, int p, float (&v)[kMaxVD]) {
#pragma unroll
    for (int c = 0; c < kMaxVD; c++) {
        float x = 0.0f;
        if (c < s.nval) x = s.val[(int64_t)c * s.val_cs + p];
        else if (c == s.nval && s.homog) x = 1.0f;
        v[c] = x;
    }
}

// b[r] for a runtime r without dynamic register indexing.
template <int D>
__device__ __forceinline__ float pick_weight(const Simplex<D> &s, int r) {
    float w = s.b[0];
#pragma unroll
    for (int q = 1; q <= D; q++) w = (q == r) ? s.b[q] : w;
    return w;
}

// ---------------------------------------------------------------------------
// clear: table words to kEmpty, accumulators and counters to zero (a fill
// kernel rather than hipMemsetAsync, so the sequence can be graph-captured).
__global__ void k_clear(uint64_t *__restrict__ tkeys, size_t nwords, float *__restrict__ tvals,
                        size_t nvals, Counters *ctr) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t0; i < nwords; i += stride) tkeys[i] = kEmpty;
    for (size_t i = t0; i < nvals; i += stride) tvals[i] = 0.0f;
    if (t0 == 0) {
        ctr->created = 0; ctr->flags = 0; ctr->listed = 0; ctr->done_pixels = 0;
        for (int i = 0; i < kMaxDH; i++) { ctr->kmin[i] = 0x7fffffff; ctr->kmax[i] = -1; }
        ctr->probes = 0; ctr->inserts = 0; ctr->atomics = 0;
    }
}

// Adds this wave's created count and statistics to the global counters.
__device__ __forceinline__ void flush_counts(Counters *ctr, unsigned created, uint32_t probes,
                                             uint32_t inserts, uint32_t atomics, bool count) {
    const unsigned c = wave_sum(created);
    if (__lane_id() == 0 && c) atomicAdd(&ctr->created, c);
    if (count) {
        const unsigned long long pr = wave_sum((unsigned long long)probes);
        const unsigned long long in = wave_sum((unsigned long long)inserts);
        const unsigned long long at = wave_sum((unsigned long long)atomics);
        if (__lane_id() == 0) {
            atomicAdd(&ctr->probes, pr);
            atomicAdd(&ctr->inserts, in);
            atomicAdd(&ctr->atomics, at);
        }
    }
}

// ---------------------------------------------------------------------------
// splat, direct form: every (pixel, vertex) inserts into the global table and
// adds weight*value with FP32 atomics (src/Permutohedral.h:365-392).
template <int D, bool BILAT>
__global__ __launch_bounds__(256) void k_splat_direct(PixSrc src, int n, int row_w, int tile_w, int vd,
                                                      LatConst lc, TableArgs ta, Counters *ctr,
                                                      int count) {
    constexpr int W = KW<D>::value;
    const int p = pixel_index(blockIdx.x, threadIdx.x, blockDim.x, n, row_w, tile_w);
    const bool stop = __hip_atomic_load(&ctr->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kFlagFull;
    bool active = (p >= 0) && !stop;
    unsigned created = 0;
    uint32_t probes = 0, inserts = 0, atomics = 0;
    if (active) {
        float pos[D], v[kMaxVD];
        load_position<D, BILAT>(src, p, pos);
        load_values(src, p, v);
        Simplex<D> s;
        simplex<D>(pos, lc.sf, s);
        if (s.bad) {
            atomicOr(&ctr->flags, kFlagBad);
        } else {
            bool ok = true;
#pragma unroll 1
            for (int r = 0; r <= D; r++) {
                uint64_t k[W];
                pack_vertex<D>(s, r, k);
                bool cr;
                const int slot = table_insert<D>(ta.tkeys, ta.mask, k, hash_key<D>(k), cr, probes);
                inserts++;
                if (slot < 0) { ok = false; break; }
                created += cr ? 1u : 0u;
                float *dst = ta.tvals + (size_t)slot * vd;
                const float w = pick_weight<D>(s, r);
#pragma unroll
                for (int c = 0; c < kMaxVD; c++)
                    if (c < vd) atomicAdd(dst + c, w * v[c]);
                atomics += vd;
            }
            if (!ok) atomicOr(&ctr->flags, kFlagFull);
        }
    }
    flush_counts(ctr, created, probes, inserts, atomics, count != 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        // count pixels of workgroups that ran to completion (regrowth estimate)
        atomicAdd(&ctr->done_pixels, stop ? 0u : (unsigned)blockDim.x);
    }
    if (__lane_id() == 0) {
        const unsigned c = __hip_atomic_load(&ctr->created, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c > ta.max_fill) atomicOr(&ctr->flags, kFlagFull);
    }
}

// splat, deduplicated form: the workgroup first inserts its (pixel, vertex)
// keys into an LDS table and pre-reduces weight*value there (LDS FP32
// atomics), then each distinct key goes to the global table once with one
// atomic add per channel.  Pixels of a tile share most simplex vertices, so
// this divides global probes and atomics by the sharing factor.
template <int D, bool BILAT, int LT>
__global__ __launch_bounds__(256) void k_splat_dedup(PixSrc src, int n, int row_w, int tile_w, int vd,
                                                     LatConst lc, TableArgs ta, Counters *ctr,
                                                     int count) {
    constexpr int W = KW<D>::value;
    constexpr int LVD = 4;   // LDS accumulator channels (dedup is used for vd <= 4)
    __shared__ uint64_t lkeys[LT * W];
    __shared__ float lvals[LT * LVD];
    for (int i = threadIdx.x; i < LT * W; i += blockDim.x) lkeys[i] = kEmpty;
    for (int i = threadIdx.x; i < LT * LVD; i += blockDim.x) lvals[i] = 0.0f;
    const int p = pixel_index(blockIdx.x, threadIdx.x, blockDim.x, n, row_w, tile_w);
    const bool stop = __hip_atomic_load(&ctr->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kFlagFull;
    __syncthreads();

    if (p >= 0 && !stop) {
        float pos[D], v[kMaxVD];
        load_position<D, BILAT>(src, p, pos);
        load_values(src, p, v);
        Simplex<D> s;
        simplex<D>(pos, lc.sf, s);
        if (s.bad) {
            atomicOr(&ctr->flags, kFlagBad);
        } else {
#pragma unroll 1
            for (int r = 0; r <= D; r++) {
                uint64_t k[W];
                pack_vertex<D>(s, r, k);
                // LDS insert-or-find, same word-claim protocol as the global
                // table.  LT exceeds the workgroup's key count (checked by the
                // launcher), so a slot is always found within LT probes.
                uint32_t ls = (uint32_t)(hash_key<D>(k) >> 40) & (LT - 1);
                for (int it = 0; it < LT; it++) {
                    bool match = true;
#pragma unroll
                    for (int w = 0; w < W; w++) {
                        if (match) {
                            uint64_t *a = &lkeys[ls * W + w];
                            uint64_t cur = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (cur == kEmpty) {
                                uint64_t e = kEmpty;
                                const bool ok = __hip_atomic_compare_exchange_strong(
                                    a, &e, k[w], __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                cur = ok ? k[w] : e;
                            }
                            match = (cur == k[w]);
                        }
                    }
                    if (match) break;
                    ls = (ls + 1) & (LT - 1);
                }
                const float w = pick_weight<D>(s, r);
#pragma unroll
                for (int c = 0; c < LVD; c++)
                    if (c < vd) atomicAdd(&lvals[ls * LVD + c], w * v[c]);
            }
        }
    }
    __syncthreads();

    // Flush the distinct keys to the global table.
    unsigned created = 0;
    uint32_t probes = 0, inserts = 0, atomics = 0;
    bool full = false;
    for (int ls = threadIdx.x; ls < LT; ls += blockDim.x) {
        if (lkeys[ls * W] == kEmpty) continue;
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = lkeys[ls * W + w];
        bool cr;
        const int slot = table_insert<D>(ta.tkeys, ta.mask, k, hash_key<D>(k), cr, probes);
        inserts++;
        if (slot < 0) { full = true; continue; }
        created += cr ? 1u : 0u;
        float *dst = ta.tvals + (size_t)slot * vd;
#pragma unroll
        for (int c = 0; c < LVD; c++)
            if (c < vd) atomicAdd(dst + c, lvals[ls * LVD + c]);
        atomics += vd;
    }
    if (full) atomicOr(&ctr->flags, kFlagFull);
    flush_counts(ctr, created, probes, inserts, atomics, count != 0);
    if (threadIdx.x == 0) atomicAdd(&ctr->done_pixels, stop ? 0u : (unsigned)blockDim.x);
    if (__lane_id() == 0) {
        const unsigned c = __hip_atomic_load(&ctr->created, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c > ta.max_fill) atomicOr(&ctr->flags, kFlagFull);
    }
}

// ---------------------------------------------------------------------------
// compaction: occupied slots -> list (wave-aggregated append), plus the
// per-coordinate min/max of the biased keys for the radix-sort bit budget.
template <int D>
__global__ __launch_bounds__(256) void k_compact(TableArgs ta, uint32_t cap, uint32_t *__restrict__ list,
                                                 Counters *ctr) {
    constexpr int W = KW<D>::value;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t k[W];
    bool occ = false;
    if (s < cap) {
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = ta.tkeys[(size_t)s * W + w];
        occ = (k[0] != kEmpty);
    } else {
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = kEmpty;
    }
    const uint64_t m = __ballot(occ);
    const int cnt = __popcll(m);
    if (cnt == 0) return;  // wave-uniform
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(&ctr->listed, (unsigned)cnt);
    base = __shfl(base, 0, kWave);
    if (occ) list[base + __popcll(m & lanemask_lt())] = s;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const int u = (int)((k[i / 4] >> (48 - 16 * (i % 4))) & 0xffff);
        int mn = occ ? u : 0x7fffffff;
        int mx = occ ? u : -1;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = min(mn, __shfl_xor(mn, off, kWave));
            mx = max(mx, __shfl_xor(mx, off, kWave));
        }
        if (__lane_id() == 0) {
            atomicMin(&ctr->kmin[i], mn);
            atomicMax(&ctr->kmax[i], mx);
        }
    }
}

// Packs (key - min) per coordinate into a `pw`-word big integer, coordinate 0
// most significant; bit offsets are precomputed on the host.
struct PackArgs {
    int umin[kMaxD];
    int off[kMaxD];
};

template <int D>
__global__ __launch_bounds__(256) void k_pack(TableArgs ta, const uint32_t *__restrict__ list, int M,
                                              PackArgs pa, int pw, uint64_t *__restrict__ pk) {
    constexpr int W = KW<D>::value;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const uint32_t s = list[i];
    uint64_t k[W];
#pragma unroll
    for (int w = 0; w < W; w++) k[w] = ta.tkeys[(size_t)s * W + w];
    uint64_t out[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < D; c++) {
        const uint64_t rel = (uint64_t)((int)((k[c / 4] >> (48 - 16 * (c % 4))) & 0xffff) - pa.umin[c]);
        const int off = pa.off[c];
        const int w0 = off >> 6, sh = off & 63;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            if (w == w0) out[w] |= rel << sh;
            if (w == w0 + 1 && sh != 0) out[w] |= rel >> (64 - sh);
        }
    }
#pragma unroll
    for (int w = 0; w < 4; w++)
        if (w < pw) pk[(size_t)w * M + i] = out[w];
}

// finalize: sorted position i -> vertex id i.  Records slot -> id, copies
// the key, and converts the FP32 splat sums to the storage type.
template <typename T>
__global__ __launch_bounds__(256) void k_finalize(TableArgs ta, int kw, const uint32_t *__restrict__ sorted,
                                                  int M, int vd, int vs, uint64_t *__restrict__ vkeys,
                                                  T *__restrict__ vals, float *__restrict__ splat_copy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const uint32_t s = sorted[i];
    ta.slot_id[s] = i;
    for (int w = 0; w < kw; w++) vkeys[(size_t)i * kw + w] = ta.tkeys[(size_t)s * kw + w];
    float o[kMaxVD];
#pragma unroll
    for (int c = 0; c < kMaxVD; c++) o[c] = (c < vd) ? ta.tvals[(size_t)s * vd + c] : 0.0f;
    switch (vs) {
    case 1: store_row<T, 1>(vals, i, o); break;
    case 2: store_row<T, 2>(vals, i, o); break;
    case 4: store_row<T, 4>(vals, i, o); break;
    case 8: store_row<T, 8>(vals, i, o); break;
    default: store_row<T, 16>(vals, i, o); break;
    }
    if (splat_copy)
        for (int c = 0; c < vd; c++) splat_copy[(size_t)i * vd + c] = o[c];
}

// Neighbour table (src/Permutohedral.h:446-451): along axis j < D the two
// neighbours are k + 1 - (D+1) e_j and k - 1 + (D+1) e_j on the first D
// coordinates; axis D moves the implicit coordinate, i.e. k + 1 and k - 1.
template <int D>
__global__ __launch_bounds__(256) void k_neighbors(TableArgs ta, const uint64_t *__restrict__ vkeys, int M,
                                                   int2 *__restrict__ nbr, Counters *ctr, int count) {
    constexpr int W = KW<D>::value;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t probes = 0;
    if (i < M) {
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = vkeys[(size_t)i * W + w];
        int c[D];
#pragma unroll
        for (int q = 0; q < D; q++) c[q] = key_coord<D>(k, q);
#pragma unroll 1
        for (int j = 0; j <= D; j++) {
            uint64_t a[W], b[W];
#pragma unroll
            for (int w = 0; w < W; w++) { a[w] = 0; b[w] = 0; }
#pragma unroll
            for (int q = 0; q < D; q++) {
                const int da = (q == j) ? -D : 1;
                const int db = (q == j) ? D : -1;
                set_key_coord<D>(a, q, c[q] + da);
                set_key_coord<D>(b, q, c[q] + db);
            }
            const int sa = table_find<D>(ta.tkeys, ta.mask, a, hash_key<D>(a), probes);
            const int sb = table_find<D>(ta.tkeys, ta.mask, b, hash_key<D>(b), probes);
            nbr[(size_t)j * M + i] = make_int2(sa >= 0 ? ta.slot_id[sa] : -1, sb >= 0 ? ta.slot_id[sb] : -1);
        }
    }
    if (count) {
        const unsigned long long pr = wave_sum((unsigned long long)probes);
        if (__lane_id() == 0) atomicAdd(&ctr->probes, pr);
    }
}

// One blur pass along one axis (src/Permutohedral.h:442-475): Jacobi, new =
// 0.25*a + 0.5*self + 0.25*b evaluated left to right, missing neighbour = 0.
template <typename T, int VS>
__global__ __launch_bounds__(256) void k_blur(const int2 *__restrict__ nbr, int M, const T *__restrict__ in,
                                              T *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int2 nb = nbr[i];
    float a[kMaxVD], o[kMaxVD], b[kMaxVD], r[kMaxVD];
#pragma unroll
    for (int c = 0; c < VS; c++) { a[c] = 0.0f; b[c] = 0.0f; }
    if (nb.x >= 0) load_row<T, VS>(in, nb.x, a);
    load_row<T, VS>(in, i, o);
    if (nb.y >= 0) load_row<T, VS>(in, nb.y, b);
#pragma unroll
    for (int c = 0; c < VS; c++) r[c] = (0.25f * a[c] + 0.5f * o[c]) + 0.25f * b[c];
    store_row<T, VS>(out, i, r);
}

// Slice (src/Permutohedral.h:405-427): recompute the simplex, find the D+1
// vertices, sum weight * value in vertex order, then write raw channels or
// normalise by the homogeneous channel (src/GaussTransform.cpp:673-676).
// A vertex absent from the lattice (slice positions differ from splat
// positions) contributes zero.
template <int D, bool BILAT, typename T>
__global__ __launch_bounds__(256) void k_slice(PixSrc src, int n, int vd, int vs, LatConst lc, TableArgs ta,
                                               const T *__restrict__ vals, OutDst dst, Counters *ctr,
                                               int count) {
    constexpr int W = KW<D>::value;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t probes = 0;
    if (p < n) {
        float pos[D];
        load_position<D, BILAT>(src, p, pos);
        Simplex<D> s;
        simplex<D>(pos, lc.sf, s);
        float acc[kMaxVD];
#pragma unroll
        for (int c = 0; c < kMaxVD; c++) acc[c] = 0.0f;
        if (s.bad) {
            atomicOr(&ctr->flags, kFlagBad);
        } else {
#pragma unroll 1
            for (int r = 0; r <= D; r++) {
                uint64_t k[W];
                pack_vertex<D>(s, r, k);
                const int slot = table_find<D>(ta.tkeys, ta.mask, k, hash_key<D>(k), probes);
                if (slot < 0) continue;
                const int id = ta.slot_id[slot];
                float v[kMaxVD];
                load_row_rt<T>(vals, id, vs, v);
                const float w = pick_weight<D>(s, r);
#pragma unroll
                for (int c = 0; c < kMaxVD; c++) acc[c] = acc[c] + w * v[c];
            }
        }
        if (dst.normalize) {
            float h = acc[0];
#pragma unroll
            for (int c = 1; c < kMaxVD; c++) h = (c == vd - 1) ? acc[c] : h;
#pragma unroll
            for (int c = 0; c < kMaxVD; c++)
                if (c < vd - 1) dst.out[(int64_t)c * dst.out_cs + p] = acc[c] / h;
        } else {
#pragma unroll
            for (int c = 0; c < kMaxVD; c++)
                if (c < vd) dst.out[(int64_t)c * dst.out_cs + p] = acc[c];
        }
    }
    if (count) {
        const unsigned long long pr = wave_sum((unsigned long long)probes);
        if (__lane_id() == 0) atomicAdd(&ctr->probes, pr);
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_unpack_keys(const uint64_t *__restrict__ vkeys, int M,
                                                     int16_t *__restrict__ keys) {
    constexpr int W = KW<D>::value;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    uint64_t k[W];
#pragma unroll
    for (int w = 0; w < W; w++) k[w] = vkeys[(size_t)i * W + w];
#pragma unroll
    for (int q = 0; q < D; q++) keys[(size_t)i * D + q] = (int16_t)key_coord<D>(k, q);
}

template <typename T>
__global__ __launch_bounds__(256) void k_store_to_f32(const T *__restrict__ vals, int M, int vd, int vs,
                                                      float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    for (int c = 0; c < vd; c++) out[(size_t)i * vd + c] = to_f32<T>(vals[(size_t)i * vs + c]);
}

// ---------------------------------------------------------------------------
// launchers

static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#define GT_DISPATCH_D(d, MACRO)                                                   \
    switch (d) {                                                                  \
    case 1: MACRO(1); break;  case 2: MACRO(2); break;  case 3: MACRO(3); break;  \
    case 4: MACRO(4); break;  case 5: MACRO(5); break;  case 6: MACRO(6); break;  \
    case 7: MACRO(7); break;  case 8: MACRO(8); break;  case 9: MACRO(9); break;  \
    case 10: MACRO(10); break; case 11: MACRO(11); break; case 12: MACRO(12); break; \
    case 13: MACRO(13); break; case 14: MACRO(14); break; case 15: MACRO(15); break; \
    case 16: MACRO(16); break;                                                    \
    default: return hipErrorInvalidValue;                                         \
    }

hipError_t launch_clear(uint64_t *tkeys, size_t nwords, float *tvals, size_t nvals, Counters *ctr,
                        hipStream_t st) {
    const size_t m = nwords > nvals ? nwords : nvals;
    unsigned g = blocks_for((int64_t)m, 256);
    if (g > 4096) g = 4096;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_clear, dim3(g), dim3(256), 0, st, tkeys, nwords, tvals, nvals, ctr);
    return hipGetLastError();
}

hipError_t launch_splat(int d, bool bilat, bool dedup, const PixSrc &src, int n, int row_w, int tile_w,
                        int block, int vd, const LatConst &lc, const TableArgs &ta, Counters *ctr,
                        bool count, hipStream_t st) {
    if (block <= 0 || block > 256) block = 256;
    block &= ~63;
    if (block == 0) block = 64;
    if (tile_w > 0 && (block % tile_w) != 0) tile_w = 0;
    int64_t nblk;
    if (tile_w > 0 && row_w > 0) {
        const int tiles_x = (row_w + tile_w - 1) / tile_w;
        const int th = block / tile_w;
        const int rows = n / row_w;
        nblk = (int64_t)tiles_x * ((rows + th - 1) / th);
    } else {
        tile_w = 0;
        nblk = blocks_for(n, block);
    }
    if (nblk == 0) return hipSuccess;
    const int cnt = count ? 1 : 0;
#define GT_SPLAT(D)                                                                                   \
    {                                                                                                 \
        if (dedup && D <= 6 && vd <= 4) {                                                             \
            constexpr int LT = 2048; /* > 256 * (D+1) keys of a workgroup */                          \
            if (bilat) hipLaunchKernelGGL((k_splat_dedup<D, true, LT>), dim3(nblk), dim3(block), 0, st, \
                                          src, n, row_w, tile_w, vd, lc, ta, ctr, cnt);               \
            else hipLaunchKernelGGL((k_splat_dedup<D, false, LT>), dim3(nblk), dim3(block), 0, st,     \
                                    src, n, row_w, tile_w, vd, lc, ta, ctr, cnt);                     \
        } else {                                                                                      \
            if (bilat) hipLaunchKernelGGL((k_splat_direct<D, true>), dim3(nblk), dim3(block), 0, st,   \
                                          src, n, row_w, tile_w, vd, lc, ta, ctr, cnt);               \
            else hipLaunchKernelGGL((k_splat_direct<D, false>), dim3(nblk), dim3(block), 0, st,        \
                                    src, n, row_w, tile_w, vd, lc, ta, ctr, cnt);                     \
        }                                                                                             \
    }
    GT_DISPATCH_D(d, GT_SPLAT)
#undef GT_SPLAT
    return hipGetLastError();
}

hipError_t launch_compact(int d, const TableArgs &ta, uint32_t capacity, uint32_t *slot_list, Counters *ctr,
                          hipStream_t st) {
#define GT_COMPACT(D) hipLaunchKernelGGL(k_compact<D>, dim3(blocks_for(capacity, 256)), dim3(256), 0, st, \
                                         ta, capacity, slot_list, ctr)
    GT_DISPATCH_D(d, GT_COMPACT)
#undef GT_COMPACT
    return hipGetLastError();
}

hipError_t launch_pack(int d, const TableArgs &ta, const uint32_t *slot_list, int M, const int *umin,
                       const int *off, int pw, uint64_t *pk, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    PackArgs pa;
    for (int i = 0; i < kMaxD; i++) { pa.umin[i] = i < d ? umin[i] : 0; pa.off[i] = i < d ? off[i] : 0; }
#define GT_PACK(D) hipLaunchKernelGGL(k_pack<D>, dim3(blocks_for(M, 256)), dim3(256), 0, st, ta, slot_list, \
                                      M, pa, pw, pk)
    GT_DISPATCH_D(d, GT_PACK)
#undef GT_PACK
    return hipGetLastError();
}

hipError_t launch_finalize(int d, int storage, const TableArgs &ta, const uint32_t *sorted, int M, int vd,
                           int vs, uint64_t *vkeys, void *vals, float *splat_copy, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const int kw = (d + 3) / 4;
    const dim3 g(blocks_for(M, 256)), b(256);
    if (storage == 0)
        hipLaunchKernelGGL(k_finalize<float>, g, b, 0, st, ta, kw, sorted, M, vd, vs, vkeys, (float *)vals, splat_copy);
    else if (storage == 1)
        hipLaunchKernelGGL(k_finalize<__hip_bfloat16>, g, b, 0, st, ta, kw, sorted, M, vd, vs, vkeys,
                           (__hip_bfloat16 *)vals, splat_copy);
    else
        hipLaunchKernelGGL(k_finalize<__half>, g, b, 0, st, ta, kw, sorted, M, vd, vs, vkeys, (__half *)vals,
                           splat_copy);
    return hipGetLastError();
}

hipError_t launch_neighbors(int d, const TableArgs &ta, const uint64_t *vkeys, int M, int2 *nbr, Counters *ctr,
                            bool count, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const int cnt = count ? 1 : 0;
#define GT_NBR(D) hipLaunchKernelGGL(k_neighbors<D>, dim3(blocks_for(M, 256)), dim3(256), 0, st, ta, vkeys, M, \
                                     nbr, ctr, cnt)
    GT_DISPATCH_D(d, GT_NBR)
#undef GT_NBR
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_blur_t(int vs, const int2 *nbr, int M, const void *in, void *out, hipStream_t st) {
    const dim3 g(blocks_for(M, 256)), b(256);
    const T *i = (const T *)in;
    T *o = (T *)out;
    switch (vs) {
    case 1: hipLaunchKernelGGL((k_blur<T, 1>), g, b, 0, st, nbr, M, i, o); break;
    case 2: hipLaunchKernelGGL((k_blur<T, 2>), g, b, 0, st, nbr, M, i, o); break;
    case 4: hipLaunchKernelGGL((k_blur<T, 4>), g, b, 0, st, nbr, M, i, o); break;
    case 8: hipLaunchKernelGGL((k_blur<T, 8>), g, b, 0, st, nbr, M, i, o); break;
    case 16: hipLaunchKernelGGL((k_blur<T, 16>), g, b, 0, st, nbr, M, i, o); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_blur(int storage, int vs, const int2 *nbr, int M, const void *in, void *out, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    if (storage == 0) return launch_blur_t<float>(vs, nbr, M, in, out, st);
    if (storage == 1) return launch_blur_t<__hip_bfloat16>(vs, nbr, M, in, out, st);
    return launch_blur_t<__half>(vs, nbr, M, in, out, st);
}

hipError_t launch_slice(int d, bool bilat, int storage, const PixSrc &src, int n, int vd, int vs,
                        const LatConst &lc, const TableArgs &ta, const void *vals, const OutDst &dst,
                        Counters *ctr, bool count, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const dim3 g(blocks_for(n, 256)), b(256);
    const int cnt = count ? 1 : 0;
#define GT_SLICE_T(D, T)                                                                                  \
    {                                                                                                     \
        if (bilat) hipLaunchKernelGGL((k_slice<D, true, T>), g, b, 0, st, src, n, vd, vs, lc, ta,        \
                                      (const T *)vals, dst, ctr, cnt);                                    \
        else hipLaunchKernelGGL((k_slice<D, false, T>), g, b, 0, st, src, n, vd, vs, lc, ta,             \
                                (const T *)vals, dst, ctr, cnt);                                          \
    }
#define GT_SLICE(D)                                            \
    {                                                          \
        if (storage == 0) GT_SLICE_T(D, float)                 \
        else if (storage == 1) GT_SLICE_T(D, __hip_bfloat16)   \
        else GT_SLICE_T(D, __half)                             \
    }
    GT_DISPATCH_D(d, GT_SLICE)
#undef GT_SLICE
#undef GT_SLICE_T
    return hipGetLastError();
}

hipError_t launch_unpack_keys(int d, const uint64_t *vkeys, int M, int16_t *keys, hipStream_t st) {
    if (M <= 0) return hipSuccess;
#define GT_UNPACK(D) hipLaunchKernelGGL(k_unpack_keys<D>, dim3(blocks_for(M, 256)), dim3(256), 0, st, vkeys, M, keys)
    GT_DISPATCH_D(d, GT_UNPACK)
#undef GT_UNPACK
    return hipGetLastError();
}

hipError_t launch_store_to_f32(int storage, const void *vals, int M, int vd, int vs, float *out, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const dim3 g(blocks_for(M, 256)), b(256);
    if (storage == 0) hipLaunchKernelGGL(k_store_to_f32<float>, g, b, 0, st, (const float *)vals, M, vd, vs, out);
    else if (storage == 1)
        hipLaunchKernelGGL(k_store_to_f32<__hip_bfloat16>, g, b, 0, st, (const __hip_bfloat16 *)vals, M, vd, vs, out);
    else hipLaunchKernelGGL(k_store_to_f32<__half>, g, b, 0, st, (const __half *)vals, M, vd, vs, out);
    return hipGetLastError();
}

} // namespace gt

namespace gt {

// Strided view -> strided view copy, used to gather the non-dense views of
// JointBilateral's zero-sigma recursion into dense scratch images and back.
__global__ __launch_bounds__(256) void k_view_copy(const float *__restrict__ src, int64_t sys, int64_t sts,
                                                   int64_t scs, float *__restrict__ dst, int64_t dys, int64_t dts,
                                                   int64_t dcs, int W, int H, int F, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = (int)(i % W);
    int64_t q = i / W;
    const int y = (int)(q % H);
    q /= H;
    const int t = (int)(q % F);
    const int64_t c = q / F;
    dst[c * dcs + t * dts + y * dys + x] = src[c * scs + t * sts + y * sys + x];
}

hipError_t launch_view_copy(const float *src, int64_t sys, int64_t sts, int64_t scs, float *dst, int64_t dys,
                            int64_t dts, int64_t dcs, int W, int H, int F, int C, hipStream_t st) {
    const int64_t total = (int64_t)W * H * F * C;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_view_copy, dim3(blocks_for(total, 256)), dim3(256), 0, st, src, sys, sts, scs, dst, dys,
                       dts, dcs, W, H, F, total);
    return hipGetLastError();
}

} // namespace gt
