// Lattice kernel templates for gfx950: position bounds, splat (elevate + hash insert +
// accumulate), compaction, key packing for the radix sort, finalize (sorted
// ids), neighbour table, blur passes and slice.  The launchers (gt_launch_*.hip)
// map the runtime lattice dimension d and key format to template parameters.
#pragma once
#include <type_traits>

#include "gt_device.h"
#include "gt_internal.h"

namespace gt {

constexpr int kBlock = 256;          // threads per workgroup of every kernel here

// ---------------------------------------------------------------------------
// pixel sources

// Pixel index of thread `tid` of workgroup `blk`: plain linear order, or a
// tile_w x (kBlock/tile_w) tile of an image with rows of row_w pixels
// (tile_w a divisor of kBlock, so a power of two).  A tile also gives the
// pixel's column x and row (t * H + y), which spares load_position its
// divisions; in linear order x = row = -1.
__device__ __forceinline__ int pixel_index(int blk, int tid, int n, int row_w, int tile_w, int &x, int &row) {
    if (tile_w <= 0 || row_w <= 0) {
        const int64_t p = (int64_t)blk * kBlock + tid;
        x = row = -1;
        return p < n ? (int)p : -1;
    }
    const int lg = __builtin_ctz((unsigned)tile_w);
    const int tiles_x = (row_w + tile_w - 1) >> lg;
    const int bx = blk % tiles_x, by = blk / tiles_x;   // the same for the whole workgroup
    x = (bx << lg) + (tid & (tile_w - 1));
    row = by * (kBlock >> lg) + (tid >> lg);             // kBlock / tile_w rows per tile
    const int rows = n / row_w;
    return (x < row_w && row < rows) ? row * row_w + x : -1;
}

__device__ __forceinline__ int pixel_index(int blk, int tid, int n, int row_w, int tile_w) {
    int x, row;
    return pixel_index(blk, tid, n, row_w, tile_w, x, row);
}

// Position of pixel p at column x and row t * H + y (x < 0: not known, found
// from p).
template <int D, bool BILAT>
__device__ __forceinline__ void load_position(const PixSrc &s, int p, int x, int row, float (&pos)[D]) {
    if constexpr (BILAT) {
        // (ref_c * (1.0f/sigma_r), x*(1.0f/sigma_x), y*(1.0f/sigma_y),
        //  t*(1.0f/sigma_t)) for the filtered axes, in that order
        // (src/GaussTransform.cpp:654-666).
        if (x < 0) {
            row = p / s.W;
            x = p - row * s.W;
        }
        int y = row, t = 0;
        if (s.F > 1) {   // (also when t is not filtered: the rows still run over the frames)
            t = row / s.H;
            y = row - t * s.H;
        }
        const float xs = (float)x * s.ix, ys = (float)y * s.iy, ts = (float)t * s.it;
        const float a0 = s.fx ? xs : (s.fy ? ys : ts);
        const float a1 = s.fx ? (s.fy ? ys : ts) : ts;
        // Every channel's load is issued before any is used (a branch per
        // channel would wait for each in turn): channels past the guide's
        // read its last plane again, a cache hit, and are not used.
        float g[D];
#pragma unroll
        for (int c = 0; c < D; c++) g[c] = 0.0f;
        if (s.cr > 0) {
#pragma unroll
            for (int c = 0; c < D; c++) g[c] = s.ref[(int64_t)min(c, s.cr - 1) * s.ref_cs + p];
        }
#pragma unroll
        for (int c = 0; c < D; c++) {
            const int k = c - s.cr;
            pos[c] = (c < s.cr) ? g[c] * s.ic : ((k == 0) ? a0 : (k == 1 ? a1 : ts));
        }
    } else {
        // pos[c] = splat(x,y,t,c) * invSigma[c] (src/GaussTransform.cpp:176-178)
#pragma unroll
        for (int c = 0; c < D; c++) pos[c] = s.pos[(int64_t)c * s.pos_cs + p] * s.inv[c];
    }
}

template <int D, bool BILAT>
__device__ __forceinline__ void load_position(const PixSrc &s, int p, float (&pos)[D]) {
    load_position<D, BILAT>(s, p, -1, -1, pos);
}

// Values of pixel p: nval planes, then a constant 1 when homog, zeros after.
// All loads are issued before any is used (channels past the planes read the
// last plane again, a cache hit).
template <int NV>
__device__ __forceinline__ void load_values(const PixSrc &s, int p, float (&v)[kMaxVD]) {
    float l[kMaxVD];
#pragma unroll
    for (int c = 0; c < NV; c++) l[c] = 0.0f;
    if (s.nval > 0) {
#pragma unroll
        for (int c = 0; c < NV; c++) l[c] = s.val[(int64_t)min(c, s.nval - 1) * s.val_cs + p];
    }
#pragma unroll
    for (int c = 0; c < NV; c++) v[c] = (c < s.nval) ? l[c] : ((c == s.nval && s.homog) ? 1.0f : 0.0f);
}

// OR of per-lane flags, one atomic per wave that has any (rare paths only).
__device__ __forceinline__ void report_flags(Counters *ctr, unsigned f) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off, kWave);
    if (__lane_id() == 0 && f) atomicOr(&ctr->flags, f);
}

__device__ __forceinline__ void report_counts(Counters *ctr, uint32_t probes, uint32_t inserts, uint32_t atomics) {
    const unsigned long long pr = wave_sum((unsigned long long)probes);
    const unsigned long long in = wave_sum((unsigned long long)inserts);
    const unsigned long long at = wave_sum((unsigned long long)atomics);
    if (__lane_id() == 0) {
        if (pr) atomicAdd(&ctr->probes, pr);
        if (in) atomicAdd(&ctr->inserts, in);
        if (at) atomicAdd(&ctr->atomics, at);
    }
}

// Lists the occupied slots of an LDS tile table (order of slot index within a
// wave, waves appended through one LDS counter).
template <int W, int LT>
__device__ __forceinline__ void list_tile(const uint64_t *lk, uint16_t *llist, unsigned *lcount) {
    static_assert(LT % kBlock == 0, "tile table size must be a multiple of the workgroup size");
    for (int base = 0; base < LT; base += kBlock) {
        const int ls = base + threadIdx.x;
        const bool occ = lk[ls * W] != kEmpty;
        const uint64_t m = __ballot(occ);
        unsigned wbase = 0;
        if (__lane_id() == 0 && m) wbase = atomicAdd(lcount, (unsigned)__popcll(m));
        wbase = __shfl(wbase, 0, kWave);
        if (occ) llist[wbase + __popcll(m & lanemask_lt())] = (uint16_t)ls;
    }
}

// ---------------------------------------------------------------------------
// Tiles.  A tiled kernel runs one workgroup per tile (a loop over tiles keeps
// the kernel arguments live across iterations: 100 VGPRs and SGPR spills
// instead of 67, occupancy 4 instead of 7).  A batch maps a global tile to its
// image by binary search over the images' first tiles; the image index
// becomes the key tag, so the images' lattices never mix.
struct TileView {
    PixSrc src;
    OutDst dst;
    int n, row_w, local;
    uint64_t tag;
};

// Workgroup -> tile order that gives each XCD a contiguous run of tiles (a
// band of the image): workgroups are dispatched round-robin over the 8 XCDs,
// so workgroup b runs on XCD b % 8.  Neighbouring tiles share most of their
// lattice vertices; in one band their table lines stay in one XCD's L2
// (2048^2 splat 190 -> 179 us).  A permutation of [0, g), so any dispatch
// order stays correct.  Batches keep the round-robin order: their tiles'
// cost varies from image to image, and bands of it unbalance the XCDs
// (64-image batch 20.0 -> 23.3 ms banded).
__device__ __forceinline__ int xcd_tile(int b, int g) {
    constexpr int kXcd = 8;
    const int per = g / kXcd, rem = g % kXcd, x = b % kXcd, k = b / kXcd;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// Vertices of workgroup b of g (one thread each, kBlock a round): ids i,
// i + step, ... below end.  Without a part map, one round of b * kBlock +
// thread.  With one, the workgroups on XCD x (b % 8, g a multiple of 8) walk
// the ids of the x-th eighth of the table parts in increasing order (part
// by part), so that their gathers stay in that XCD's L2.
__device__ __forceinline__ void part_walk(const PartMap &pm, unsigned M, unsigned b, unsigned g, unsigned &i,
                                          unsigned &end, unsigned &step) {
    if (pm.cpp == 0) {
        i = b * kBlock + threadIdx.x;
        end = M;
        step = g * kBlock;
        return;
    }
    // (after a full table M is 0 while the chunk offsets are not: clamp)
    const unsigned x = b & 7u;
    i = pm.off[x * pm.cpp] + (b >> 3) * kBlock + threadIdx.x;
    end = min(x == 7 ? M : pm.off[(x + 1) * pm.cpp], M);
    step = (g >> 3) * kBlock;
}

template <bool BATCH>
__device__ __forceinline__ void tile_view(int tile, const PixSrc &src0, const OutDst &dst0, int n0, int row_w0,
                                          const BatchArgs &ba, const KeyPlan &kp, TileView &tv) {
    tv.src = src0;
    tv.dst = dst0;
    if constexpr (!BATCH) {
        tv.n = n0;
        tv.row_w = row_w0;
        tv.local = tile;
        tv.tag = 0;
    } else {
        int lo = 0, hi = ba.count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ba.img[mid].tile0 <= tile) lo = mid;
            else hi = mid - 1;
        }
        const BatchImage &bi = ba.img[lo];
        const int64_t plane = (int64_t)bi.W * bi.H;
        tv.src.ref = bi.ref;
        tv.src.ref_cs = plane;
        tv.src.W = bi.W;
        tv.src.H = bi.H;
        tv.src.ic = bi.ic;
        tv.src.ix = bi.ix;
        tv.src.iy = bi.iy;
        tv.src.val = bi.im;
        tv.src.val_cs = plane;
        tv.dst.out = bi.out;
        tv.dst.out_cs = plane;
        tv.n = (int)plane;
        tv.row_w = bi.W;
        tv.local = tile - bi.tile0;
        tv.tag = (uint64_t)lo << kp.coord_bits;
    }
}

// ---------------------------------------------------------------------------
// splat, tiled form: the workgroup inserts its 256 pixels' D+1 vertex keys
// into an LDS tile table and pre-reduces weight*value there (LDS FP32
// atomics); then each distinct key is inserted into the global table once and
// its sums added with one FP32 atomic per channel.  Pixels of a tile share
// most simplex vertices, so global probes and atomics drop by that factor.
// (src/Permutohedral.h:365-392)
//
// LDS accumulators are 64-bit fixed point (ValScale), channel-major
// (lv[c*LT + slot]): LDS integer atomics run over ten times faster than LDS
// FP32 atomics on gfx950, and the sums are exact.  LT = 512 slots keep the
// workgroup's LDS small enough for high occupancy; keys that do not find an
// LDS slot within kLdsProbes go to the global table directly.
constexpr int kLdsProbes = 32;

// Rotates a[0..N-1] left by rot (0 <= rot < N) with log2(N) stages of
// compile-time-indexed selects (no dynamic register indexing).
template <int N, typename T>
__device__ __forceinline__ void rotate_left(T (&a)[N], int rot) {
#pragma unroll
    for (int sh = 1; sh < N; sh <<= 1) {
        const bool on = (rot & sh) != 0;
        T t[N];
#pragma unroll
        for (int i = 0; i < N; i++) t[i] = a[(i + sh) % N];
#pragma unroll
        for (int i = 0; i < N; i++) a[i] = on ? t[i] : a[i];
    }
}

// Static LDS of the tiled kernels (for the waves-per-CU padding).
template <int W, int LVD, int LT>
constexpr size_t splat_tile_lds() { return (size_t)LT * W * 8 + (size_t)LVD * LT * 8 + (size_t)LT * 2 + 16; }
template <int W, int LVD, int LT>
constexpr size_t slice_tile_lds() { return (size_t)LT * W * 8 + (size_t)LT * LVD * 4 + (size_t)LT * 2 + 16; }

// NaN/Inf values (src/Permutohedral.h:377 adds barycentric * value in FP32,
// so they spread through the sums): the fixed-point sums take such channels
// as 0, and each of the pixel's vertices gets a mark per channel instead --
// the class of the FP32 product weight * value: +Inf (bit 2c), -Inf (bit
// 2c+1) or NaN (both bits; a +Inf and a -Inf term also sum to NaN).  The
// finalize turns a marked sum into NaN, +Inf or -Inf.  Only calls whose value
// bounds saw a NaN/Inf take this path.
template <int D, bool PACKED, bool BATCH>
__device__ __forceinline__ void mark_nonfinite(const Simplex<D> &s, const KeyPlan &kp, uint64_t tag, const TableArgs &ta,
                                            const float (&v)[kMaxVD], unsigned nfmask, int vd, uint32_t &probes,
                                            unsigned &flags) {
    using K = KeyOps<D, PACKED>;
    constexpr int W = K::W;
    for (int r = 0; r <= D; r++) {
        uint64_t k[W];
        if (!K::make(s, r, kp, tag, k)) continue;   // reported by the caller
        const int slot = table_insert<W>(ta.tkeys, ta.mask, k, home_slot<D, PACKED, BATCH>(ta, kp, k), probes);
        if (slot < 0) { flags |= insert_fail_flag(slot); continue; }
        write_tag<W>(ta, slot, k);
        uint32_t bits = 0;
        for (int c = 0; c < vd; c++) {
            if (!((nfmask >> c) & 1u)) continue;
            const float prod = s.b[r] * v[c];
            const uint32_t cls = (prod != prod) ? 3u : (prod > 0.0f ? 1u : 2u);
            bits |= cls << (2 * c);
        }
        if (bits) atomicOr(ta.tnf + slot, bits);
    }
}

// Splits a pixel's values into what the fixed-point sums take (NaN/Inf as 0)
// and a mask of the non-finite channels (the originals stay in v).
template <int NV>
__device__ __forceinline__ unsigned split_nonfinite(const float (&v)[kMaxVD], int vd, float (&fin)[kMaxVD]) {
    unsigned m = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) {
        const bool bad = c < vd && !(fabsf(v[c]) <= 3.402823466e38f);
        m |= bad ? (1u << c) : 0u;
        fin[c] = bad ? 0.0f : v[c];
    }
    return m;
}

template <int D, bool PACKED, bool BILAT, int LVD, int LT, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_splat_tile(PixSrc src0, int n0, int row_w0, int tile_w, BatchArgs ba,
                                                       int vd, LatConst lc, KeyPlan kp,
                                                       const ValScale *__restrict__ vsc, TableArgs ta,
                                                       Replay rp, Counters *ctr, int count) {
    using K = KeyOps<D, PACKED>;
    constexpr int W = K::W;
    __shared__ uint64_t lk[LT * W];
    __shared__ unsigned long long lv[LVD * LT];
    __shared__ uint16_t llist[LT];
    __shared__ unsigned lcount, lfull;
    unsigned flags = 0;
    uint32_t probes = 0, inserts = 0, atomics = 0;
    const OutDst no_dst{nullptr, 0, 0};
    TileView tv;
    const int tile = BATCH ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
    tile_view<BATCH>(tile, src0, no_dst, n0, row_w0, ba, kp, tv);
    const unsigned rq = (unsigned)tile * kBlock + threadIdx.x;   // this pixel's replay entry
    const PixSrc &src = tv.src;
    for (int i = threadIdx.x; i < LT * W; i += kBlock) lk[i] = kEmpty;
    for (int i = threadIdx.x; i < LVD * LT; i += kBlock) lv[i] = 0ull;
    if (threadIdx.x == 0) {
        lcount = 0;
        lfull = *(volatile const unsigned *)&ctr->flags & kFlagFull;
    }
    int px, prow;
    const int p = pixel_index(tv.local, threadIdx.x, tv.n, tv.row_w, tile_w, px, prow);
    __syncthreads();
    // One reading of the full flag for the whole workgroup: other workgroups
    // may set it at any time, and the flush's second phase reads the slots
    // its first phase parked, so every thread must skip it or none.
    const bool full = lfull != 0;

    float sc[LVD];
#pragma unroll
    for (int c = 0; c < LVD; c++) sc[c] = vsc->s[c];
    const bool nfcall = vsc->nonfinite != 0;

    if (p >= 0) {
        float pos[D], v[kMaxVD], vorig[kMaxVD];
        load_values<LVD>(src, p, v);   // (issued first: its loads and the positions' are in flight together)
        load_position<D, BILAT>(src, p, px, prow, pos);
        unsigned nfmask = 0;
        if (nfcall) {
#pragma unroll
            for (int c = 0; c < LVD; c++) vorig[c] = v[c];
            nfmask = split_nonfinite<LVD>(vorig, vd, v);
        }
        // values in fixed-point units once per pixel: w * (v s) rounds as (w v) s
#pragma unroll
        for (int c = 0; c < LVD; c++) v[c] *= sc[c];
        Simplex<D> s;
        simplex<D>(pos, lc.sf, s);
        if (s.bad) {
            flags |= kFlagBad;
        } else {
            if (nfmask && !full) mark_nonfinite<D, PACKED, BATCH>(s, kp, tv.tag, ta, vorig, nfmask, vd, probes, flags);
            // Lane i visits the vertices starting at rank i mod (D+1): lanes of a
            // wave that share a key (neighbouring pixels, same rank) update it in
            // different instructions, which removes most LDS atomic conflicts.
            uint64_t key[D + 1][W];
            float wt[D + 1];
            int okr[D + 1];
            if constexpr (PACKED) {
                uint64_t kk[D + 1];
                const int ok = packed_vertex_keys<D>(s, kp, tv.tag, kk) ? 1 : 0;
#pragma unroll
                for (int r = 0; r <= D; r++) {
                    key[r][0] = kk[r];
                    okr[r] = ok;
                }
            } else {
#pragma unroll
                for (int r = 0; r <= D; r++) okr[r] = K::make(s, r, kp, tv.tag, key[r]) ? 1 : 0;
            }
#pragma unroll
            for (int r = 0; r <= D; r++) wt[r] = s.b[r];
            const int rot = (int)(threadIdx.x % (D + 1));
            if constexpr (!PACKED) rotate_left<D + 1>(okr, rot);   // packed: one flag for all vertices
            rotate_left<D + 1>(wt, rot);
#pragma unroll
            for (int w = 0; w < W; w++) {
                uint64_t col[D + 1];
#pragma unroll
                for (int r = 0; r <= D; r++) col[r] = key[r][w];
                rotate_left<D + 1>(col, rot);
#pragma unroll
                for (int r = 0; r <= D; r++) key[r][w] = col[r];
            }
            constexpr int RW = replay_dwords(D);
            uint32_t rslot[RW];   // the replay entry
#pragma unroll
            for (int i = 0; i < RW; i++) rslot[i] = 0u;
#pragma unroll
            for (int r = 0; r <= D; r++) {
                const uint64_t(&k)[W] = key[r];
                if (!okr[r]) {   // (the call is replanned and redone)
                    flags |= kFlagRange;
                    rslot[r / 3] |= kReplayGlobal << (10 * (r % 3));
                    continue;
                }
                const int ls = lds_insert<W, LT, kLdsProbes>(lk, k);
                // (in the rotated order; the slice undoes the rotation)
                rslot[r / 3] |= (ls >= 0 ? (uint32_t)ls : kReplayGlobal) << (10 * (r % 3));
                if (ls >= 0) {
#pragma unroll
                    for (int c = 0; c < LVD; c++)
                        if (c < vd) atomicAdd(&lv[c * LT + ls], fixed_word(wt[r] * v[c]));
                } else if (!full) {
                    const int slot = table_insert<W>(ta.tkeys, ta.mask, k, home_slot<D, PACKED, BATCH>(ta, kp, k), probes);
                    inserts++;
                    if (rp.rec) rp.gs[r * rp.stride + rq] = (uint32_t)slot;
                    if (slot < 0) { flags |= insert_fail_flag(slot); continue; }
                    write_tag<W>(ta, slot, k);
                    unsigned long long *dst = ta.tacc + (size_t)slot * vd;
#pragma unroll
                    for (int c = 0; c < LVD; c++)
                        if (c < vd) atomicAdd(dst + c, fixed_word(wt[r] * v[c]));
                    atomics += vd;
                }
            }
            if (rp.rec) {   // one vector store, streamed past the L2 that holds the table
                if constexpr (RW == 2) {
                    const u32x2 v = {rslot[0], rslot[1]};
                    __builtin_nontemporal_store(v, reinterpret_cast<u32x2 *>(rp.rec) + rq);
                } else {
                    const u32x4 v = {rslot[0], rslot[1], rslot[2], rslot[3]};
                    __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(rp.rec) + rq);
                }
            }
        }
    }
    __syncthreads();
    list_tile<W, LT>(lk, llist, &lcount);
    __syncthreads();

    // Flush, once any insert has found the table full skipping the global
    // work (the call is redone with a larger table; a stale read only costs
    // time).  First one thread per distinct key inserts it into the global
    // table, all keys of the tile in one round of table latency, and parks the
    // slot in the key's tile-table word (-1: table full).  Then LVD lanes per
    // key, lane c adding channel c, so one instruction's atomics cover whole
    // contiguous vertex rows; those need no reply and nothing waits on them.
    const unsigned cnt = full ? 0u : lcount;
    for (unsigned j = threadIdx.x; j < cnt; j += kBlock) {
        const int ls = (int)llist[j];
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = lk[ls * W + w];
        const int slot = table_insert<W>(ta.tkeys, ta.mask, k, home_slot<D, PACKED, BATCH>(ta, kp, k), probes);
        inserts++;
        if (slot < 0) flags |= insert_fail_flag(slot);
        write_tag<W>(ta, slot, k);
        lk[ls * W] = (uint64_t)(int64_t)slot;
        if (rp.rec) rp.tent[(size_t)tile * LT + j] = make_uint2((unsigned)ls, (unsigned)slot);
    }
    if (rp.rec && threadIdx.x == 0) rp.tcnt[tile] = cnt;
    __syncthreads();
    const int c = (int)(threadIdx.x % LVD);
    for (unsigned base = 0; base < cnt; base += kBlock / LVD) {
        const unsigned j = base + threadIdx.x / LVD;
        if (j < cnt && c < vd) {
            const int ls = (int)llist[j];
            const int slot = (int)(int64_t)lk[ls * W];
            if (slot >= 0) {
                atomicAdd(ta.tacc + (size_t)slot * vd + c, lv[c * LT + ls]);
                atomics++;
            }
        }
    }
    report_flags(ctr, flags);
    if (count) report_counts(ctr, probes, inserts, atomics);
}

// splat, direct form (D > 10, vd > 4, or the LDS tiles turned off): every
// (pixel, vertex) goes to the global table, its values added as fixed-point
// integers.  Workgroups take the tiles in XCD bands, as the tiled form does.
template <int D, bool PACKED, bool BILAT>
__global__ __launch_bounds__(kBlock) void k_splat_direct(PixSrc src, int n, int row_w, int tile_w, int vd,
                                                         LatConst lc, KeyPlan kp,
                                                         const ValScale *__restrict__ vsc, TableArgs ta,
                                                         Counters *ctr, int count) {
    using K = KeyOps<D, PACKED>;
    constexpr int W = K::W;
    int px, prow;
    const int p = pixel_index(xcd_tile(blockIdx.x, gridDim.x), threadIdx.x, n, row_w, tile_w, px, prow);
    unsigned flags = 0;
    uint32_t probes = 0, inserts = 0, atomics = 0;
    const bool full = *(volatile const unsigned *)&ctr->flags & kFlagFull;
    if (p >= 0 && !full) {
        float pos[D], v[kMaxVD], vorig[kMaxVD];
        load_values<kMaxVD>(src, p, v);   // (issued first: its loads and the positions' are in flight together)
        load_position<D, BILAT>(src, p, px, prow, pos);
        unsigned nfmask = 0;
        if (vsc->nonfinite) {
#pragma unroll
            for (int c = 0; c < kMaxVD; c++) vorig[c] = v[c];
            nfmask = split_nonfinite<kMaxVD>(vorig, vd, v);
        }
        // values in fixed-point units once per pixel: w * (v s) rounds as (w v) s
#pragma unroll
        for (int c = 0; c < kMaxVD; c++)
            if (c < vd) v[c] *= vsc->s[c];
        Simplex<D> s;
        simplex<D>(pos, lc.sf, s);
        if (s.bad) {
            flags |= kFlagBad;
        } else {
            if (nfmask) mark_nonfinite<D, PACKED, false>(s, kp, 0ull, ta, vorig, nfmask, vd, probes, flags);
            uint64_t pk[D + 1];
            bool pk_ok = true;
            if constexpr (PACKED) pk_ok = packed_vertex_keys<D>(s, kp, 0ull, pk);
#pragma unroll
            for (int r = 0; r <= D; r++) {
                uint64_t k[W];
                if constexpr (PACKED) {
                    k[0] = pk[r];
                    if (!pk_ok) { flags |= kFlagRange; continue; }
                } else {
                    if (!K::make(s, r, kp, 0ull, k)) { flags |= kFlagRange; continue; }
                }
                const int slot = table_insert<W>(ta.tkeys, ta.mask, k, home_slot<D, PACKED, false>(ta, kp, k), probes);
                inserts++;
                if (slot < 0) { flags |= insert_fail_flag(slot); continue; }
                write_tag<W>(ta, slot, k);
                unsigned long long *dst = ta.tacc + (size_t)slot * vd;
#pragma unroll
                for (int c = 0; c < kMaxVD; c++)
                    if (c < vd) atomicAdd(dst + c, fixed_word(s.b[r] * v[c]));
                atomics += vd;
            }
        }
    }
    report_flags(ctr, flags);
    if (count) report_counts(ctr, probes, inserts, atomics);
}

// ---------------------------------------------------------------------------
// compaction: occupied slots -> list, in slot order, by count / scan / write.
template <int W>
__global__ __launch_bounds__(kBlock) void k_count(const uint64_t *__restrict__ tkeys, uint32_t cap,
                                                  uint32_t *__restrict__ counts, const uint8_t *__restrict__ tags) {
    __shared__ unsigned ws[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kCompactSlots;
    unsigned c = 0;
    if (tags) {   // occupancy from the tags: one byte per slot instead of a key line
#pragma unroll
        for (int k = 0; k < kCompactSlots / (4 * kBlock); k++) {
            const size_t s = base + 4 * ((size_t)k * kBlock + threadIdx.x);
            if (s < cap) {
                const uint32_t t = *reinterpret_cast<const uint32_t *>(tags + s);
                c += (unsigned)__popc(t & 0x80808080u);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kCompactSlots / kBlock; k++) {
            const size_t s = base + (size_t)k * kBlock + threadIdx.x;
            c += (s < cap && tkeys[s * W] != kEmpty) ? 1u : 0u;
        }
    }
    c = wave_sum(c);
    if (__lane_id() == 0) ws[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < kBlock / kWave; w++) t += ws[w];
        counts[blockIdx.x] = t;
    }
}

template <int W, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_write_list(const uint64_t *__restrict__ tkeys, uint32_t cap,
                                                       const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ list, uint64_t *__restrict__ sort_keys) {
    constexpr int NW = kBlock / kWave;
    __shared__ unsigned ws[NW];
    const size_t base = (size_t)blockIdx.x * kCompactSlots;
    unsigned run = offsets[blockIdx.x];
    const int wave = threadIdx.x / kWave;
    for (int k = 0; k < kCompactSlots / kBlock; k++) {
        const size_t s = base + (size_t)k * kBlock + threadIdx.x;
        uint64_t w0 = kEmpty;
        if (s < cap) w0 = tkeys[s * W];
        const bool occ = w0 != kEmpty;
        const uint64_t m = __ballot(occ);
        if (__lane_id() == 0) ws[wave] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned off = run, tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            off += (w < wave) ? ws[w] : 0u;
            tot += ws[w];
        }
        if (occ) {
            const unsigned pos = off + (unsigned)__popcll(m & lanemask_lt());
            list[pos] = (uint32_t)s;
            if (PACKED) sort_keys[pos] = w0;
        }
        run += tot;
        __syncthreads();
    }
}

// Compaction and finalize in one pass, for ids in compaction order (no
// sort): occupied slot s gets id pos, its rank among the occupied slots (as
// k_write_list ranks them), and k_finalize's work for it is done here: the
// key copied to vkeys[pos], the sums converted to vals[pos] by the same
// expression and set back to zero, slot -> id recorded (in slot_id, or
// folded into the key word).  After a full table (M = 0) nothing is written.
// The sums of slot s in FP32 (one rounding: int64 -> double is exact below
// 2^53, then one FP32 rounding), NaN/Inf where the slot holds non-finite
// marks; the sums and marks read are set back to zero.
__device__ __forceinline__ void take_sums(const TableArgs &ta, const ValScale *__restrict__ vsc, size_t s, int vd,
                                          float (&o)[kMaxVD]) {
#pragma unroll
    for (int c = 0; c < kMaxVD; c++)
        o[c] = (c < vd) ? (float)((double)(long long)ta.tacc[s * vd + c] * (double)vsc->inv_s[c]) : 0.0f;
    for (int c = 0; c < vd; c++) ta.tacc[s * vd + c] = 0ull;
    if (vsc->nonfinite) {
        const uint32_t m = ta.tnf[s];
        if (m) {
#pragma unroll
            for (int c = 0; c < kMaxVD; c++) {
                const uint32_t b = (m >> (2 * c)) & 3u;
                o[c] = b == 3u ? __builtin_nanf("") : b == 1u ? INFINITY : b == 2u ? -INFINITY : o[c];
            }
            ta.tnf[s] = 0u;
        }
    }
}

// take_sums with the sums already in registers (vd <= 4).
__device__ __forceinline__ void take_sums4(const TableArgs &ta, const ValScale *__restrict__ vsc, size_t s, int vd,
                                           const unsigned long long (&a)[4], float (&o)[kMaxVD]) {
#pragma unroll
    for (int c = 0; c < kMaxVD; c++)
        o[c] = (c < 4 && c < vd) ? (float)((double)(long long)a[c < 4 ? c : 0] * (double)vsc->inv_s[c]) : 0.0f;
    for (int c = 0; c < vd; c++) ta.tacc[s * vd + c] = 0ull;
    if (vsc->nonfinite) {
        const uint32_t m = ta.tnf[s];
        if (m) {
#pragma unroll
            for (int c = 0; c < kMaxVD; c++) {
                const uint32_t b = (m >> (2 * c)) & 3u;
                o[c] = b == 3u ? __builtin_nanf("") : b == 1u ? INFINITY : b == 2u ? -INFINITY : o[c];
            }
            ta.tnf[s] = 0u;
        }
    }
}

template <int W, typename T>
__global__ __launch_bounds__(kBlock) void k_write_final(TableArgs ta, uint32_t cap,
                                                        const uint32_t *__restrict__ offsets, const Counters *ctr,
                                                        const ValScale *__restrict__ vsc, int vd, int vs,
                                                        uint64_t *__restrict__ vkeys, T *__restrict__ vals,
                                                        unsigned *__restrict__ track, SelfScan sc, NbrInit ninit) {
    constexpr int NW = kBlock / kWave;
    const int wave = threadIdx.x / kWave;
    const int chunk = (int)blockIdx.x;
    unsigned run;
    if (sc.offs) {
        // offsets holds the chunks' counts: this chunk's offset and the total
        __shared__ unsigned spre[NW], stot[NW];
        unsigned pre = 0, tot = 0;
        {   // all of a thread's counts in flight at once (nchunks <= kSelfScanChunks)
            constexpr int U = kSelfScanChunks / kBlock;
            unsigned v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = (int)threadIdx.x + u * kBlock;
                v[u] = i < sc.nchunks ? offsets[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = (int)threadIdx.x + u * kBlock;
                tot += v[u];
                pre += i < chunk ? v[u] : 0u;
            }
        }
        pre = wave_sum(pre);
        tot = wave_sum(tot);
        if (__lane_id() == 0) {
            spre[wave] = pre;
            stot[wave] = tot;
        }
        __syncthreads();
        pre = tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            pre += spre[w];
            tot += stot[w];
        }
        const bool over = tot > sc.limit;
        const bool full = over || (*(volatile const unsigned *)&ctr->flags & kFlagFull);   // the same everywhere
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (over) atomicOr(&sc.wctr->flags, kFlagFull);
            sc.wctr->M = full ? 0u : tot;
            sc.wctr->filled = tot;
        }
        if (threadIdx.x == 0) sc.offs[chunk] = pre;
        if (full) return;
        run = pre;
    } else {
        if (*(volatile const unsigned *)&ctr->flags & kFlagFull) return;   // the same for every thread
        run = offsets[chunk];
    }
    float mx[kMaxVD];   // track: largest finite magnitude per channel (a narrow storage's first scale)
#pragma unroll
    for (int c = 0; c < kMaxVD; c++) mx[c] = 0.0f;
    const unsigned run0 = run;   // this chunk's ids are [run0, run at the end)
    const size_t base = (size_t)chunk * kCompactSlots;
    // Each thread's KI key words first (all in flight at once) and one
    // barrier for the ranks of all of them; then row by row, with the next
    // row's sums (vd <= 4) loaded before this row's stores, so the chunk's
    // memory round trips overlap instead of following one another.
    constexpr int KI = kCompactSlots / kBlock;
    __shared__ unsigned wsk[KI][NW];
    uint64_t w0[KI], mk[KI];
#pragma unroll
    for (int k = 0; k < KI; k++) {   // (with tags: an occupied slot's tag is set, so only those keys are read)
        const size_t s = base + (size_t)k * kBlock + threadIdx.x;
        w0[k] = (s < cap && (!ta.tags || ta.tags[s] != 0)) ? ta.tkeys[s * W] : kEmpty;
    }
#pragma unroll
    for (int k = 0; k < KI; k++) {
        mk[k] = __ballot(w0[k] != kEmpty);
        if (__lane_id() == 0) wsk[k][wave] = (unsigned)__popcll(mk[k]);
    }
    __syncthreads();
    const bool pre = vd <= 4;
    unsigned long long acc[2][4];   // row k's sums, and row k+1's in flight
    auto load_sums = [&](int k, unsigned long long(&a)[4]) {
        const size_t s = base + (size_t)k * kBlock + threadIdx.x;
#pragma unroll
        for (int c = 0; c < 4; c++) a[c] = (pre && w0[k] != kEmpty && c < vd) ? ta.tacc[s * vd + c] : 0ull;
    };
    load_sums(0, acc[0]);
#pragma unroll
    for (int k = 0; k < KI; k++) {
        if (k + 1 < KI) load_sums(k + 1, acc[(k + 1) & 1]);
        const size_t s = base + (size_t)k * kBlock + threadIdx.x;
        unsigned off = run, tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            off += (w < wave) ? wsk[k][w] : 0u;
            tot += wsk[k][w];
        }
        if (w0[k] != kEmpty) {
            const unsigned pos = off + (unsigned)__popcll(mk[k] & lanemask_lt());
            uint64_t kw[W];
            kw[0] = w0[k];
#pragma unroll
            for (int w = 1; w < W; w++) kw[w] = ta.tkeys[s * W + w];
#pragma unroll
            for (int w = 0; w < W; w++) vkeys[(size_t)pos * W + w] = kw[w];

            float o[kMaxVD];
            if (pre) take_sums4(ta, vsc, s, vd, acc[k & 1], o);
            else take_sums(ta, vsc, s, vd, o);
            store_row_rt<T>(vals, pos, vs, o);
            if (track) take_max<kMaxVD>(mx, o);
            if (W == 1 && ta.idshift > 0) ta.tkeys[s] = w0[k] | ((uint64_t)pos << ta.idshift);
            else ta.slot_id[s] = (int)pos;
        }
        run += tot;
    }
    if (ninit.nbr && ninit.grid == 0) {   // this chunk's neighbour entries, {-1, -1} on every axis
        const u32x2 none = {~0u, ~0u};
        for (int j = 0; j < ninit.axes; j++)
            for (unsigned i = run0 + threadIdx.x; i < run; i += kBlock)
                __builtin_nontemporal_store(none, reinterpret_cast<u32x2 *>(ninit.nbr) + (size_t)j * ninit.stride + i);
    }
    if (track) publish_max<kMaxVD>(track, mx, vd);
}

// Wide keys -> packed sort key (big integer of pw words, word 0 least
// significant) by the plan's offsets; coordinates outside the plan flag a
// replan.
template <int D>
__global__ __launch_bounds__(kBlock) void k_pack(TableArgs ta, const uint32_t *__restrict__ list, const Counters *ctr,
                                                 int stride, KeyPlan kp, int pw, uint64_t *__restrict__ pk,
                                                 Counters *wctr) {
    using K = KeyOps<D, false>;
    constexpr int W = K::W;
    const unsigned M = ctr->M;
    const unsigned i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M) return;
    const uint32_t s = list[i];
    uint64_t k[W];
#pragma unroll
    for (int w = 0; w < W; w++) k[w] = ta.tkeys[(size_t)s * W + w];
    uint64_t out[4] = {0, 0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int c = 0; c < D; c++) {
        const uint32_t rel = (uint32_t)(K::coord(k, c, kp) - kp.lo[c]);
        ok &= rel <= kp.span[c];
        const uint64_t v = rel & kp.span[c];
        const int off = kp.shift[c];
        const int w0 = off >> 6, sh = off & 63;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            if (w == w0) out[w] |= v << sh;
            if (w == w0 + 1 && sh != 0) out[w] |= v >> (64 - sh);
        }
    }
    if (!ok) atomicOr(&wctr->flags, kFlagRange);
#pragma unroll
    for (int w = 0; w < 4; w++)
        if (w < pw) pk[(size_t)w * stride + i] = out[w];
}

// finalize: position i of the sorted (or, without a sort, the compaction)
// order -> vertex id i.  Records slot -> id (in slot_id, or folded into the
// table word above the key when ta.idshift > 0), copies the key, and
// converts the fixed-point splat sums to FP32 rows (one rounding: int64 ->
// double is exact below 2^53, then one FP32 rounding; a BF16/FP16 storage
// takes its first rounding in blur pass 0, scaled by the largest magnitude
// tracked here).  The sums it has read are set back to zero, so after a
// complete call the whole sum array is zero again and the next call need not
// clear it.
template <typename T>
__device__ __forceinline__ void finalize_one(TableArgs &ta, const ValScale *__restrict__ vsc, int kw,
                                             const uint32_t *__restrict__ sorted,
                                             const uint64_t *__restrict__ sorted_keys, int vd, int vs,
                                             uint64_t *__restrict__ vkeys, T *__restrict__ vals,
                                             float *__restrict__ splat_copy, unsigned i, float (&mx)[kMaxVD],
                                             bool track) {
    const uint32_t s = sorted[i];
    if (sorted_keys) {   // packed keys: the sort's own output, read in order
        const uint64_t key = sorted_keys[i];
        vkeys[i] = key;
        if (ta.idshift > 0) ta.tkeys[s] = key | ((uint64_t)i << ta.idshift);
    } else {
        for (int w = 0; w < kw; w++) vkeys[(size_t)i * kw + w] = ta.tkeys[(size_t)s * kw + w];
        if (ta.idshift > 0) ta.tkeys[s] |= (uint64_t)i << ta.idshift;   // kw == 1
    }
    if (ta.idshift == 0) ta.slot_id[s] = (int)i;
    float o[kMaxVD];
    take_sums(ta, vsc, s, vd, o);
    store_row_rt<T>(vals, i, vs, o);
    if (track) take_max<kMaxVD>(mx, o);
    if (splat_copy)
        for (int c = 0; c < vd; c++) splat_copy[(size_t)i * vd + c] = o[c];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_finalize(TableArgs ta, const ValScale *__restrict__ vsc, int kw,
                                                     const uint32_t *__restrict__ sorted,
                                                     const uint64_t *__restrict__ sorted_keys, const Counters *ctr,
                                                     int vd, int vs, uint64_t *__restrict__ vkeys,
                                                     T *__restrict__ vals, float *__restrict__ splat_copy,
                                                     unsigned *__restrict__ track) {
    const unsigned M = ctr->M;
    const unsigned i = blockIdx.x * kBlock + threadIdx.x;
    float mx[kMaxVD];   // track: largest finite magnitude per channel (a narrow storage's first scale)
#pragma unroll
    for (int c = 0; c < kMaxVD; c++) mx[c] = 0.0f;
    if (i < M) finalize_one(ta, vsc, kw, sorted, sorted_keys, vd, vs, vkeys, vals, splat_copy, i, mx, track != nullptr);
    if (track) publish_max<kMaxVD>(track, mx, vd);
}


// Neighbour table from the sorted keys, without the hash table.  Moving a
// key by +-(1,...,1) -+ (D+1) e_j moves every field of its words by a
// constant, so when no field leaves its range the neighbour's words are the
// vertex's words plus constant deltas (exact: no carry or borrow crosses a
// field).  The targets of a run of sorted vertices are then sorted too, and
// finding them is a merge with the sorted keys.  Workgroup = 2048 consecutive
// vertices x one direction (blockIdx.y = 2 axis + sign): wave 0 finds the
// first target's lower bound by a 64-ary search, then the workgroup stages
// windows of sorted keys in LDS and each thread binary-searches its targets
// there.  Ids are sorted ranks, so the table equals a hash-built one entry
// for entry.  Packed keys use the plan's fields; wide keys (W words of four
// int16+32768 fields, coordinate 0 most significant) their 16-bit fields.
constexpr int kNbrItems = 8;
constexpr int kNbrBlock = kBlock * kNbrItems;
constexpr int kNbrWindowWords = 4096;

template <int W>
__device__ __forceinline__ bool key_less(const uint64_t *a, const uint64_t (&b)[W]) {
#pragma unroll
    for (int w = 0; w < W; w++)
        if (a[w] != b[w]) return a[w] < b[w];
    return false;
}

template <int W>
__device__ __forceinline__ bool key_eq(const uint64_t *a, const uint64_t (&b)[W]) {
    bool e = true;
#pragma unroll
    for (int w = 0; w < W; w++) e = e && a[w] == b[w];
    return e;
}

// First index in s[0, n) (W words per key) whose key is >= x (n if none), by
// one wave: 64 probes per round narrow the range 64-fold.
template <int W>
__device__ __forceinline__ unsigned wave_lower_bound(const uint64_t *__restrict__ s, unsigned n,
                                                     const uint64_t (&x)[W]) {
    const int lane = __lane_id();
    unsigned lo = 0, hi = n;   // answer in [lo, hi]
    while (hi - lo > kWave) {
        const unsigned step = (hi - lo + kWave - 1) / kWave;
        const unsigned p = lo + (unsigned)(lane + 1) * step - 1;
        const bool below = p < hi && key_less<W>(s + (size_t)p * W, x);
        const int c = __popcll(__ballot(below));   // probes below x form a prefix
        const unsigned nlo = c == 0 ? lo : lo + (unsigned)c * step;
        const unsigned cp = lo + (unsigned)(c + 1) * step - 1;
        hi = (c < kWave && cp < hi) ? cp : hi;
        lo = nlo;
    }
    const unsigned p = lo + (unsigned)lane;
    const bool below = p < hi && key_less<W>(s + (size_t)p * W, x);
    return lo + (unsigned)__popcll(__ballot(below));
}

template <int D, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_nbr_merge(const uint64_t *__restrict__ sk, KeyPlan kp,
                                                      const Counters *ctr, int stride, int *__restrict__ nbr) {
    constexpr int W = KeyOps<D, PACKED>::W;
    constexpr int WIN = kNbrWindowWords / W;   // keys per LDS window
    __shared__ uint64_t win[WIN * W];
    __shared__ uint64_t red[kBlock / kWave][W];
    __shared__ unsigned s_cur;
    const unsigned M = ctr->M;
    const unsigned b0 = blockIdx.x * kNbrBlock;
    if (b0 >= M) return;
    const int j = blockIdx.y >> 1, sgn = blockIdx.y & 1;   // sgn 0: k + 1 - (D+1)e_j, 1: k - 1 + (D+1)e_j

    // per-word deltas of the move
    uint64_t delta[W];
#pragma unroll
    for (int w = 0; w < W; w++) delta[w] = 0;
#pragma unroll
    for (int q = 0; q < D; q++) {
        const int64_t dq = (q == j) ? -(int64_t)D : 1;
        const int64_t v = sgn ? -dq : dq;
        if (PACKED) delta[0] += (uint64_t)v << kp.shift[q];
        else delta[q / 4] += (uint64_t)v << (48 - 16 * (q % 4));
    }

    uint64_t tgt[kNbrItems][W];
    bool pend[kNbrItems];
    int res[kNbrItems];
    uint64_t tmin[W];
#pragma unroll
    for (int w = 0; w < W; w++) tmin[w] = ~0ull;
#pragma unroll
    for (int r = 0; r < kNbrItems; r++) {
        const unsigned i = b0 + r * kBlock + threadIdx.x;
        bool ok = i < M;
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = ok ? sk[(size_t)i * W + w] : 0ull;
#pragma unroll
        for (int q = 0; q < D; q++) {
            uint32_t f, lo, hi;
            if (PACKED) {
                f = (uint32_t)((k[0] >> kp.shift[q]) & kp.span[q]);
                lo = 0;
                hi = kp.span[q];
            } else {
                f = (uint32_t)((k[q / 4] >> (48 - 16 * (q % 4))) & 0xffff);
                lo = 1 + 2 * (D + 1);            // the int16 key range of from_coords
                hi = 65535 - 2 * (D + 1);
            }
            const bool up = (q == j) == (sgn != 0);   // this field grows
            const uint32_t step = (q == j) ? (uint32_t)D : 1u;
            ok = ok && (up ? f + step <= hi : f >= lo + step);
        }
#pragma unroll
        for (int w = 0; w < W; w++) tgt[r][w] = k[w] + delta[w];   // exact when every field stays in range
        pend[r] = ok;
        res[r] = -1;
        if (ok && key_less<W>(tgt[r], tmin)) {
#pragma unroll
            for (int w = 0; w < W; w++) tmin[w] = tgt[r][w];
        }
    }
    // workgroup minimum of the valid targets
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        uint64_t o[W];
#pragma unroll
        for (int w = 0; w < W; w++) o[w] = __shfl_xor(tmin[w], off, kWave);
        if (key_less<W>(o, tmin)) {
#pragma unroll
            for (int w = 0; w < W; w++) tmin[w] = o[w];
        }
    }
    if (__lane_id() == 0) {
#pragma unroll
        for (int w = 0; w < W; w++) red[threadIdx.x / kWave][w] = tmin[w];
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; w++) tmin[w] = red[0][w];
    for (int v = 1; v < kBlock / kWave; v++)
        if (key_less<W>(red[v], tmin)) {
#pragma unroll
            for (int w = 0; w < W; w++) tmin[w] = red[v][w];
        }
    bool any = false;
#pragma unroll
    for (int w = 0; w < W; w++) any = any || tmin[w] != ~0ull;
    if (any) {
        if (threadIdx.x < kWave) {
            const unsigned lb = wave_lower_bound<W>(sk, M, tmin);
            if (threadIdx.x == 0) s_cur = lb;
        }
        __syncthreads();
        // cur is the same in every thread; each round either resolves every
        // pending target (the last window) or advances cur by a full window.
        // cur == M (every target beyond the largest key): nothing to find.
        unsigned cur = s_cur;
        while (cur < M) {
            const unsigned wlen = min((unsigned)WIN, M - cur);
            for (unsigned t = threadIdx.x; t < wlen * W; t += kBlock) win[t] = sk[(size_t)cur * W + t];
            __syncthreads();
            const bool last_window = cur + wlen >= M;
            bool more = false;
#pragma unroll
            for (int r = 0; r < kNbrItems; r++) {
                if (!pend[r]) continue;
                if (!last_window && key_less<W>(win + (size_t)(wlen - 1) * W, tgt[r]) ) {
                    more = true;   // beyond this window's last key
                    continue;
                }
                unsigned a = 0, b = wlen;   // lower bound in win[0, wlen)
                while (a < b) {
                    const unsigned m = (a + b) >> 1;
                    if (key_less<W>(win + (size_t)m * W, tgt[r])) a = m + 1;
                    else b = m;
                }
                if (a < wlen && key_eq<W>(win + (size_t)a * W, tgt[r])) res[r] = (int)(cur + a);
                pend[r] = false;
            }
            if (!__syncthreads_or(more ? 1 : 0)) break;
            cur += wlen;   // every pending target exceeds this window's last key
        }
    }
#pragma unroll
    for (int r = 0; r < kNbrItems; r++) {
        const unsigned i = b0 + r * kBlock + threadIdx.x;
        if (i < M) nbr[2 * ((size_t)j * stride + i) + sgn] = res[r];
    }
}

// Neighbour table by hash lookups, for lattices whose vertex ids are the
// compaction order (no sort): the same +-(1,...,1) -+ (D+1) e_j moves as
// k_nbr_merge, each target looked up in the complete global table and its
// slot mapped to the vertex id.  One thread per (vertex, axis), both signs.
// Used where the lattice is sparse (high d: most neighbours are absent, and
// a miss costs one or two probes) and sorting many wide keys would dominate.
template <int D, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_nbr_hash(const uint64_t *__restrict__ vk, KeyPlan kp, const Counters *ctr,
                                                     int stride, TableArgs ta, PartMap pm, int *__restrict__ nbr) {
    constexpr int W = KeyOps<D, PACKED>::W;
    const unsigned M = ctr->M;
    unsigned i, end, step;
    part_walk(pm, M, blockIdx.x, gridDim.x, i, end, step);
    const int j = blockIdx.y;
    for (; i < end; i += step) {
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = vk[(size_t)i * W + w];
        // the two targets: sgn 0: k + 1 - (D+1)e_j, sgn 1: k - 1 + (D+1)e_j
        uint64_t t[2][W];
        bool ok[2] = {true, true};
#pragma unroll
        for (int sgn = 0; sgn < 2; sgn++)
#pragma unroll
            for (int w = 0; w < W; w++) t[sgn][w] = k[w];
#pragma unroll
        for (int q = 0; q < D; q++) {
            uint32_t f, lo, hi;
            if (PACKED) {
                f = (uint32_t)((k[0] >> kp.shift[q]) & kp.span[q]);
                lo = 0;
                hi = kp.span[q];
            } else {
                f = (uint32_t)((k[q / 4] >> (48 - 16 * (q % 4))) & 0xffff);
                lo = 1 + 2 * (D + 1);
                hi = 65535 - 2 * (D + 1);
            }
            const uint32_t mv = (q == j) ? (uint32_t)D : 1u;
            const int64_t dq = (q == j) ? -(int64_t)D : 1;
#pragma unroll
            for (int sgn = 0; sgn < 2; sgn++) {
                const bool up = (q == j) == (sgn != 0);   // this field grows
                ok[sgn] = ok[sgn] && (up ? f + mv <= hi : f >= lo + mv);
                const int64_t v = sgn ? -dq : dq;
                if (PACKED) t[sgn][0] += (uint64_t)v << kp.shift[q];
                else t[sgn][q / 4] += (uint64_t)v << (48 - 16 * (q % 4));
            }
        }
        // Both lookups probe in lockstep, so their loads are in flight together.
        // With ids folded into the table words (one-word keys), a hit carries its
        // vertex id and slot_id is not read.
        const bool fold = W == 1 && ta.idshift > 0;
        const uint64_t kmask = fold ? ta.keymask : ~0ull;
        uint32_t s[2];
        bool done[2];
        int slot[2] = {-1, -1};
        uint64_t hit[2] = {0, 0};
#pragma unroll
        for (int sgn = 0; sgn < 2; sgn++) {
            s[sgn] = home_slot<D, PACKED>(ta, kp, t[sgn]);
            done[sgn] = !ok[sgn];
        }
        for (uint32_t it = 0; it < kMaxProbes && !(done[0] && done[1]); it++) {
            uint64_t w0[2];
#pragma unroll
            for (int sgn = 0; sgn < 2; sgn++) w0[sgn] = done[sgn] ? kEmpty : ta.tkeys[(size_t)s[sgn] * W];
#pragma unroll
            for (int sgn = 0; sgn < 2; sgn++) {
                if (done[sgn]) continue;
                if (w0[sgn] == kEmpty) {
                    done[sgn] = true;
                    continue;
                }
                bool match = (w0[sgn] & kmask) == t[sgn][0];
#pragma unroll
                for (int w = 1; w < W; w++) match = match && ta.tkeys[(size_t)s[sgn] * W + w] == t[sgn][w];
                if (match) {
                    slot[sgn] = (int)s[sgn];
                    hit[sgn] = w0[sgn];
                    done[sgn] = true;
                } else {
                    s[sgn] = (s[sgn] + 1) & ta.mask;
                }
            }
        }
        int res[2];
#pragma unroll
        for (int sgn = 0; sgn < 2; sgn++)
            res[sgn] = slot[sgn] < 0 ? -1 : fold ? (int)(hit[sgn] >> ta.idshift) : ta.slot_id[slot[sgn]];
        int2 *out = (int2 *)nbr + (size_t)j * stride + i;
        *out = make_int2(res[0], res[1]);
    }
}

// The same table, one thread per vertex for all D+1 axes (used from D = 7
// and for large lattices, where the per-axis form spends most of its time
// rebuilding the key): the
// fields, their range checks and their sum are taken once, and each target
// is the key plus one constant word per sign -- +1 on every field, then
// -(D+1) (sign 0) or +(D+1) (sign 1) on field j -- as integer arithmetic,
// exact whenever no field leaves its range (the checks).  The fields of a
// target sum to the vertex's sum -1 / +1 (j < D) or +D / -D (j = D).
// HALF: only sign 0 is looked up.  u = id(k + 1 - (D+1) e_j) has this vertex
// as its sign-1 neighbour on axis j (u - 1 + (D+1) e_j = k), so the thread
// writes entry i's first half and entry u's second half; every second half
// without a writer keeps the -1 the finalize put there.
template <int D, bool PACKED, bool HALF>
__global__ __launch_bounds__(kBlock) void k_nbr_hash_vertex(const uint64_t *__restrict__ vk, KeyPlan kp,
                                                            const Counters *ctr, int stride, TableArgs ta, PartMap pm,
                                                            int *__restrict__ nbr) {
    constexpr int W = KeyOps<D, PACKED>::W;
    constexpr int NS = HALF ? 1 : 2;   // signs looked up
    const unsigned M = ctr->M;
    unsigned i, end, step;
    part_walk(pm, M, blockIdx.x, gridDim.x, i, end, step);
    const bool fold = W == 1 && ta.idshift > 0;
    const uint64_t kmask = fold ? ta.keymask : ~0ull;
    for (; i < end; i += step) {
        uint64_t k[W], one[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            // (streamed: the table's tags and key lines are what should stay in L2)
            k[w] = __builtin_nontemporal_load(vk + (size_t)i * W + w);
            one[w] = 0;
        }
        uint32_t no_up1 = 0, no_dn1 = 0, ok_upD = 0, ok_dnD = 0;   // bit q: field q
        int sum = 0;
        const uint32_t tag = PACKED ? key_tag(ta, kp, k[0]) : 0u;   // neighbours share it
#pragma unroll
        for (int q = 0; q < D; q++) {
            const int wq = PACKED ? 0 : q / 4;
            const int sh = PACKED ? kp.shift[q] : 48 - 16 * (q % 4);
            const uint32_t f = (uint32_t)((k[wq] >> sh) & (PACKED ? (uint64_t)kp.span[q] : 0xffffull));
            const uint32_t lo = PACKED ? 0u : (uint32_t)(1 + 2 * (D + 1));
            const uint32_t hi = PACKED ? kp.span[q] : (uint32_t)(65535 - 2 * (D + 1));
            one[wq] += 1ull << sh;
            sum += (int)f;
            if (f + 1 > hi) no_up1 |= 1u << q;
            if (f < lo + 1) no_dn1 |= 1u << q;
            if (f + D <= hi) ok_upD |= 1u << q;
            if (f >= lo + D) ok_dnD |= 1u << q;
        }
        // Targets: every word moves by +-one (the same for all axes), then
        // axis j < D moves word j/4 (packed: the one word) by -+big.  Their
        // hashes are the sum of the word hashes of k +- one, with word j/4's
        // replaced.
        uint64_t gw[2][W], gsum[2] = {0, 0};
#pragma unroll
        for (int sgn = 0; sgn < NS; sgn++) {
            gsum[sgn] = 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
                gw[sgn][w] = hash_word(sgn ? k[w] - one[w] : k[w] + one[w], w);
                gsum[sgn] += gw[sgn][w];
            }
        }
        for (int j = 0; j <= D; j++) {
            uint64_t t[2][W];
            bool ok[2];
            int tsum[2];
            uint64_t hsum[2] = {gsum[0], gsum[1]};
#pragma unroll
            for (int w = 0; w < W; w++) {
                t[0][w] = k[w] + one[w];
                t[1][w] = k[w] - one[w];
            }
            if (j < D) {
                const uint32_t bit = 1u << j;
                ok[0] = (no_up1 & ~bit) == 0 && (ok_dnD & bit) != 0;
                ok[1] = (no_dn1 & ~bit) == 0 && (ok_upD & bit) != 0;
                const int wj = PACKED ? 0 : j / 4;
                const int shj = PACKED ? kp.shift[j] : 48 - 16 * (j % 4);
                const uint64_t big = (uint64_t)(D + 1) << shj;
                uint64_t m0 = t[0][0], m1 = t[1][0], g0 = gw[0][0], g1 = gw[1][0];
#pragma unroll
                for (int w = 1; w < W; w++)
                    if (w == wj) {
                        m0 = t[0][w];
                        m1 = t[1][w];
                        g0 = gw[0][w];
                        g1 = gw[1][w];
                    }
                m0 -= big;
                m1 += big;
#pragma unroll
                for (int w = 0; w < W; w++)
                    if (w == wj) {
                        t[0][w] = m0;
                        t[1][w] = m1;
                    }
                hsum[0] += hash_word(m0, wj) - g0;   // one word rehashed
                if (!HALF) hsum[1] += hash_word(m1, wj) - g1;
                tsum[0] = sum - 1;
                tsum[1] = sum + 1;
            } else {
                ok[0] = no_up1 == 0;
                ok[1] = no_dn1 == 0;
                tsum[0] = sum + D;
                tsum[1] = sum - D;
            }
            uint32_t s[2] = {0, 0};
            bool done[2] = {true, true};   // (sign 1 stays done under HALF)
            int slot[2] = {-1, -1};
            uint64_t hit[2] = {0, 0};
            uint8_t want[2] = {0, 0};
#pragma unroll
            for (int sgn = 0; sgn < NS; sgn++) {
                const uint64_t h64 = hash_final(hsum[sgn]);   // = hash_key<W>(t[sgn])
                const uint32_t h = (uint32_t)h64;
                s[sgn] = ta.part_mul ? slot_in_part(ta, h, tsum[sgn], tag) : (h & ta.mask);
                want[sgn] = key_tag8(h64);
                done[sgn] = !ok[sgn];
            }
            // through the tags: key words only where a tag matches
            if (ta.tags) {
#pragma unroll
                for (int sgn = 0; sgn < NS; sgn++)
                    if (!done[sgn]) slot[sgn] = tag_find<W>(ta, t[sgn], kmask, s[sgn], want[sgn], hit[sgn]);
            }
            for (uint32_t it = 0; !ta.tags && it < kMaxProbes && !(done[0] && done[1]); it++) {
                uint64_t w0[2];
#pragma unroll
                for (int sgn = 0; sgn < NS; sgn++) w0[sgn] = done[sgn] ? kEmpty : ta.tkeys[(size_t)s[sgn] * W];
#pragma unroll
                for (int sgn = 0; sgn < NS; sgn++) {
                    if (done[sgn]) continue;
                    if (w0[sgn] == kEmpty) {
                        done[sgn] = true;
                        continue;
                    }
                    bool match = (w0[sgn] & kmask) == t[sgn][0];
#pragma unroll
                    for (int w = 1; w < W; w++) match = match && ta.tkeys[(size_t)s[sgn] * W + w] == t[sgn][w];
                    if (match) {
                        slot[sgn] = (int)s[sgn];
                        hit[sgn] = w0[sgn];
                        done[sgn] = true;
                    } else {
                        s[sgn] = (s[sgn] + 1) & ta.mask;
                    }
                }
            }
            int res[2] = {-1, -1};
#pragma unroll
            for (int sgn = 0; sgn < NS; sgn++)
                res[sgn] = slot[sgn] < 0 ? -1 : fold ? (int)(hit[sgn] >> ta.idshift) : ta.slot_id[slot[sgn]];
            if (HALF) {
                nbr[2 * ((size_t)j * stride + i)] = res[0];
                if (res[0] >= 0) nbr[2 * ((size_t)j * stride + (unsigned)res[0]) + 1] = (int)i;
            } else {
                const u32x2 pair = {(unsigned)res[0], (unsigned)res[1]};
                __builtin_nontemporal_store(pair, reinterpret_cast<u32x2 *>(nbr) + (size_t)j * stride + i);
            }
        }
    }
}

// One blur pass along one axis (src/Permutohedral.h:442-475): Jacobi, new =
// 0.25*a + 0.5*self + 0.25*b evaluated left to right, missing neighbour = 0.
// SCALED (BF16/FP16 storage): the input holds true * 2^e_in (in_bits: its
// scale source; none for the FP32 output of the finalize), the output is
// written as true * 2^e_out (out_bits: the largest magnitude of this pass's
// input), and the output's own largest finite magnitude goes to out_max for
// the next pass.  The FP32 arithmetic is on the true values; the power-of-two
// factors are exact.
template <typename TI, typename TO, int VS, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_blur(const int2 *__restrict__ nbr, const Counters *ctr, PartMap pm,
                                                 const TI *__restrict__ in, TO *__restrict__ out,
                                                 const unsigned *__restrict__ in_bits,
                                                 const unsigned *__restrict__ out_bits, unsigned *__restrict__ out_max) {
    const unsigned M = ctr->M;
    unsigned i, end, step;
    part_walk(pm, M, blockIdx.x, gridDim.x, i, end, step);
    float ui[kMaxVD], so[kMaxVD], mx[kMaxVD];
#pragma unroll
    for (int c = 0; c < VS; c++) {
        ui[c] = (SCALED && in_bits) ? ldexpf(1.0f, -store_exp(in_bits[c])) : 1.0f;
        so[c] = SCALED ? ldexpf(1.0f, store_exp(out_bits[c])) : 1.0f;
        mx[c] = 0.0f;
    }
    for (; i < end; i += step) {
        const int2 nb = nbr[i];
        float a[kMaxVD], o[kMaxVD], b[kMaxVD], r[kMaxVD];
#pragma unroll
        for (int c = 0; c < VS; c++) { a[c] = 0.0f; b[c] = 0.0f; }
        if (nb.x >= 0) load_row<TI, VS>(in, nb.x, a);
        load_row<TI, VS>(in, i, o);
        if (nb.y >= 0) load_row<TI, VS>(in, nb.y, b);
        if (SCALED) {
#pragma unroll
            for (int c = 0; c < VS; c++) { a[c] *= ui[c]; o[c] *= ui[c]; b[c] *= ui[c]; }
        }
#pragma unroll
        for (int c = 0; c < VS; c++) r[c] = (0.25f * a[c] + 0.5f * o[c]) + 0.25f * b[c];
        if (SCALED) {
            take_max<VS>(mx, r);
#pragma unroll
            for (int c = 0; c < VS; c++) r[c] *= so[c];
        }
        store_row<TO, VS>(out, i, r);
    }
    if (SCALED) publish_max<VS>(out_max, mx, VS);
}

// Two blur passes, along axes j and k = j+1, in one launch (FP32 storage):
// out(v) = (0.25 P(u) + 0.5 P(v)) + 0.25 P(w), u and w v's neighbours on
// axis k, with P(x) = (0.25 in(a) + 0.5 in(x)) + 0.25 in(b), a and b x's
// neighbours on axis j, a missing vertex contributing 0 -- the FP32
// operations of the two passes one after the other, without the array
// between them (each P is formed three times, by the threads of v's row).
template <int VS>
__global__ __launch_bounds__(kBlock) void k_blur2(const int2 *__restrict__ nbr_j, const int2 *__restrict__ nbr_k,
                                                  const Counters *ctr, PartMap pm, const float *__restrict__ in,
                                                  float *__restrict__ out) {
    const unsigned M = ctr->M;
    unsigned i, end, step;
    part_walk(pm, M, blockIdx.x, gridDim.x, i, end, step);
    for (; i < end; i += step) {
        const int2 nk = nbr_k[i];
        const int xs[3] = {nk.x, (int)i, nk.y};
        int2 nj[3];
#pragma unroll
        for (int q = 0; q < 3; q++) nj[q] = xs[q] >= 0 ? nbr_j[xs[q]] : make_int2(-1, -1);
        float P[3][kMaxVD];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            float a[kMaxVD], o[kMaxVD], b[kMaxVD];
#pragma unroll
            for (int c = 0; c < VS; c++) { a[c] = 0.0f; o[c] = 0.0f; b[c] = 0.0f; }
            if (xs[q] >= 0) {
                if (nj[q].x >= 0) load_row<float, VS>(in, nj[q].x, a);
                load_row<float, VS>(in, xs[q], o);
                if (nj[q].y >= 0) load_row<float, VS>(in, nj[q].y, b);
            }
#pragma unroll
            for (int c = 0; c < VS; c++) P[q][c] = xs[q] >= 0 ? (0.25f * a[c] + 0.5f * o[c]) + 0.25f * b[c] : 0.0f;
        }
        float r[kMaxVD];
#pragma unroll
        for (int c = 0; c < VS; c++) r[c] = (0.25f * P[0][c] + 0.5f * P[1][c]) + 0.25f * P[2][c];
        store_row<float, VS>(out, i, r);
    }
}

// Final write of one pixel: raw vd channels, or the first vd-1 divided by the
// homogeneous last one (true division, src/GaussTransform.cpp:673-676).
template <int NV>
__device__ __forceinline__ void write_pixel(const OutDst &dst, int p, int vd, const float (&acc)[kMaxVD]) {
    if (dst.normalize) {
        float h = acc[0];
#pragma unroll
        for (int c = 1; c < NV; c++) h = (c == vd - 1) ? acc[c] : h;
#pragma unroll
        for (int c = 0; c < NV; c++)
            if (c < vd - 1) dst.out[(int64_t)c * dst.out_cs + p] = acc[c] / h;
    } else {
#pragma unroll
        for (int c = 0; c < NV; c++)
            if (c < vd) dst.out[(int64_t)c * dst.out_cs + p] = acc[c];
    }
}

// 2^-e of each channel of a narrow storage's last blur pass (sbits: its
// scale source), 1 for FP32 storage.
template <typename T, int NV>
__device__ __forceinline__ void slice_unscale(const unsigned *__restrict__ sbits, float (&u)[kMaxVD]) {
#pragma unroll
    for (int c = 0; c < NV; c++)
        u[c] = (std::is_same<T, float>::value || !sbits) ? 1.0f : ldexpf(1.0f, -store_exp(sbits[c]));
}

// Slice, tiled form (src/Permutohedral.h:405-427): the tile's distinct
// vertices are looked up once and their blurred values staged in LDS
// (channel-major); each pixel then sums weight * value over its D+1 vertices
// in vertex order.  A vertex absent from the lattice (slice positions differ
// from the splat positions) contributes nothing; a key that found no LDS
// slot is looked up in the global table directly.
template <int D, bool PACKED, bool BILAT, typename T, int LVD, int LT, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_slice_tile(PixSrc src0, int n0, int row_w0, int tile_w, BatchArgs ba,
                                                       int vd, LatConst lc, KeyPlan kp, TableArgs ta,
                                                       const T *__restrict__ vals,
                                                       const unsigned *__restrict__ sbits, OutDst dst0,
                                                       Counters *ctr, int count) {
    using K = KeyOps<D, PACKED>;
    constexpr int W = K::W;
    constexpr bool kScaled = !std::is_same<T, float>::value;
    __shared__ uint64_t lk[LT * W];
    __shared__ float lv[LT * LVD];
    __shared__ uint16_t llist[LT];
    __shared__ unsigned lcount;
    // a full table means the call is redone with a larger one: skip the slice
    if (*(volatile const unsigned *)&ctr->flags & kFlagFull) return;
    unsigned flags = 0;
    uint32_t probes = 0;
    TileView tv;
    tile_view<BATCH>(BATCH ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x), src0, dst0, n0, row_w0, ba, kp, tv);
    const PixSrc &src = tv.src;
    for (int i = threadIdx.x; i < LT * W; i += kBlock) lk[i] = kEmpty;
    if (threadIdx.x == 0) lcount = 0;
    int px, prow;
    const int p = pixel_index(tv.local, threadIdx.x, tv.n, tv.row_w, tile_w, px, prow);
    __syncthreads();

    int lsl[D + 1];   // LDS slot, -1 absent (outside the plan), -2 look up globally
    Simplex<D> s;
    bool ok = false;
#pragma unroll
    for (int r = 0; r <= D; r++) lsl[r] = -1;
    if (p >= 0) {
        float pos[D];
        load_position<D, BILAT>(src, p, px, prow, pos);
        simplex<D>(pos, lc.sf, s);
        ok = !s.bad;
        if (!ok) {
            flags |= kFlagBad;
        } else {
            // Insert in a per-lane rotated vertex order (fewer LDS conflicts,
            // as in the splat), then restore vertex order for the sums.
            uint64_t key[D + 1][W];
            int okr[D + 1];
            bool all = false;
            if constexpr (PACKED) {
                uint64_t kk[D + 1];
                all = packed_vertex_keys<D>(s, kp, tv.tag, kk);
#pragma unroll
                for (int r = 0; r <= D; r++) {
                    key[r][0] = kk[r];
                    okr[r] = 1;
                }
            }
            if (!all) {   // wide keys, or a simplex across the plan's edge (slice positions of their own)
#pragma unroll
                for (int r = 0; r <= D; r++) okr[r] = K::make(s, r, kp, tv.tag, key[r]) ? 1 : 0;
            }
            const int rot = (int)(threadIdx.x % (D + 1));
            rotate_left<D + 1>(okr, rot);
#pragma unroll
            for (int w = 0; w < W; w++) {
                uint64_t col[D + 1];
#pragma unroll
                for (int r = 0; r <= D; r++) col[r] = key[r][w];
                rotate_left<D + 1>(col, rot);
#pragma unroll
                for (int r = 0; r <= D; r++) key[r][w] = col[r];
            }
#pragma unroll
            for (int r = 0; r <= D; r++) {
                if (okr[r]) {
                    const int ls = lds_insert<W, LT, kLdsProbes>(lk, key[r]);
                    lsl[r] = ls >= 0 ? ls : -2;
                }
            }
            rotate_left<D + 1>(lsl, (D + 1 - rot) % (D + 1));
        }
    }
    __syncthreads();
    list_tile<W, LT>(lk, llist, &lcount);
    __syncthreads();

    const unsigned cnt = lcount;
    float un[kMaxVD];
    slice_unscale<T, LVD>(sbits, un);
    for (unsigned j = threadIdx.x; j < cnt; j += kBlock) {
        const int ls = llist[j];
        uint64_t k[W];
#pragma unroll
        for (int q = 0; q < W; q++) k[q] = lk[ls * W + q];
        const int id = table_vertex<D, PACKED, BATCH>(ta, kp, k, probes);
        float row[kMaxVD];
#pragma unroll
        for (int c = 0; c < LVD; c++) row[c] = 0.0f;
        if (id >= 0) load_row<T, LVD>(vals, id, row);
#pragma unroll
        for (int c = 0; c < LVD; c++) lv[c * LT + ls] = kScaled ? row[c] * un[c] : row[c];
    }
    __syncthreads();

    if (ok) {
        float acc[kMaxVD];
#pragma unroll
        for (int c = 0; c < LVD; c++) acc[c] = 0.0f;
#pragma unroll
        for (int r = 0; r <= D; r++) {
            if (lsl[r] >= 0) {
#pragma unroll
                for (int c = 0; c < LVD; c++) acc[c] = acc[c] + s.b[r] * lv[c * LT + lsl[r]];
            } else if (lsl[r] == -2) {
                uint64_t k[W];
                K::make(s, r, kp, tv.tag, k);
                const int id = table_vertex<D, PACKED, BATCH>(ta, kp, k, probes);
                if (id >= 0) {
                    float row[kMaxVD];
                    load_row<T, LVD>(vals, id, row);
#pragma unroll
                    for (int c = 0; c < LVD; c++) acc[c] = acc[c] + s.b[r] * (row[c] * un[c]);
                }
            }
        }
        write_pixel<LVD>(tv.dst, p, vd, acc);
    }
    report_flags(ctr, flags);
    if (count) report_counts(ctr, probes, 0, 0);
}

// Vertex id of a slot of the finalized table (ids folded into the key word,
// or slot_id).
__device__ __forceinline__ int slot_vertex(const TableArgs &ta, uint32_t slot) {
    return ta.idshift > 0 ? (int)(ta.tkeys[slot] >> ta.idshift) : ta.slot_id[slot];
}

// Slice at the splat's own positions from the tiled splat's replay record
// (src/Permutohedral.h:405-427): the workgroup of tile t stages the blurred
// values of the tile's distinct vertices in LDS at the splat's tile-table
// slots (their global slots give the ids, no probing); each pixel recomputes
// its simplex for the weights only (no keys, no tile-table inserts) and sums
// weight * value over its D+1 vertices in vertex order -- the same FP32
// operations as the slice that finds its vertices by key, so the output is
// the same bit for bit.  Records are checked against the table and LDS
// bounds (a call whose splat flagged an error writes records only partly and
// is discarded, but its slice still runs).
template <int D, bool BILAT, typename T, int LVD, int LT, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_slice_replay(PixSrc src0, int n0, int row_w0, int tile_w, BatchArgs ba,
                                                         int vd, LatConst lc, KeyPlan kp, TableArgs ta, Replay rp,
                                                         const T *__restrict__ vals,
                                                         const unsigned *__restrict__ sbits, OutDst dst0,
                                                         Counters *ctr, Counters *hctr) {
    constexpr bool kScaled = !std::is_same<T, float>::value;
    constexpr int RW = replay_dwords(D);
    __shared__ float lv[LT * LVD];
    // The counters are final here (this kernel changes none): hand them to
    // the host's pinned copy, which spares the call a copy launch.
    if (hctr && blockIdx.x == 0 && threadIdx.x == 0) {
        hctr->M = ctr->M;
        hctr->flags = ctr->flags;
        hctr->filled = ctr->filled;
        hctr->probes = ctr->probes;
        hctr->inserts = ctr->inserts;
        hctr->atomics = ctr->atomics;
    }
    if (*(volatile const unsigned *)&ctr->flags & kFlagFull) return;   // redone with a larger table
    const unsigned M = ctr->M;
    const int tile = BATCH ? (int)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
    TileView tv;
    tile_view<BATCH>(tile, src0, dst0, n0, row_w0, ba, kp, tv);
    float un[kMaxVD];
    slice_unscale<T, LVD>(sbits, un);
    int px, prow;
    const int p = pixel_index(tv.local, threadIdx.x, tv.n, tv.row_w, tile_w, px, prow);
    const unsigned rq = (unsigned)tile * kBlock + threadIdx.x;
    uint32_t e[RW];
    float pos[D];
    if (p >= 0) {   // (issued before the staging: the loads overlap it)
        if constexpr (RW == 2) {
            const u32x2 u = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(rp.rec) + rq);
            e[0] = u.x;
            e[1] = u.y;
        } else {
            const u32x4 u = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(rp.rec) + rq);
            e[0] = u.x;
            e[1] = u.y;
            e[2] = u.z;
            e[3] = u.w;
        }
        load_position<D, BILAT>(tv.src, p, px, prow, pos);
    }
    const unsigned cnt = min(rp.tcnt[tile], (unsigned)LT);
    {   // a thread's entries (LT / kBlock) step by step together: three
        // rounds of dependent loads (entry, id, row) instead of three each
        constexpr int SU = (LT + kBlock - 1) / kBlock;
        uint2 te[SU];
        int id[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const unsigned j = threadIdx.x + (unsigned)(u * kBlock);
            te[u] = j < cnt ? rp.tent[(size_t)tile * LT + j] : make_uint2((unsigned)LT, 0u);
        }
#pragma unroll
        for (int u = 0; u < SU; u++)
            id[u] = (te[u].x < (unsigned)LT && te[u].y <= ta.mask) ? slot_vertex(ta, te[u].y) : -1;
        float row[SU][kMaxVD];
#pragma unroll
        for (int u = 0; u < SU; u++) {
#pragma unroll
            for (int c = 0; c < LVD; c++) row[u][c] = 0.0f;
            if (id[u] >= 0 && (unsigned)id[u] < M) load_row<T, LVD>(vals, id[u], row[u]);
        }
#pragma unroll
        for (int u = 0; u < SU; u++) {
            if (te[u].x >= (unsigned)LT) continue;
#pragma unroll
            for (int c = 0; c < LVD; c++) lv[c * LT + te[u].x] = kScaled ? row[u][c] * un[c] : row[u][c];
        }
    }
    Simplex<D> s;
    if (p >= 0) simplex<D>(pos, lc.sf, s);   // (before the barrier: overlaps the staging)
    __syncthreads();
    if (p < 0 || s.bad) return;   // (a bad position fails the call)
    int ls[D + 1];
#pragma unroll
    for (int j = 0; j <= D; j++) ls[j] = (int)((e[j / 3] >> (10 * (j % 3))) & 1023u);
    // vertex order: the splat stored vertex (j + rot) % (D+1) at j
    const int rot = (int)(threadIdx.x % (D + 1));
    rotate_left<D + 1>(ls, (D + 1 - rot) % (D + 1));
    float acc[kMaxVD];
#pragma unroll
    for (int c = 0; c < LVD; c++) acc[c] = 0.0f;
#pragma unroll
    for (int r = 0; r <= D; r++) {
        if (ls[r] < LT) {
#pragma unroll
            for (int c = 0; c < LVD; c++) acc[c] = acc[c] + s.b[r] * lv[c * LT + ls[r]];
        } else if (ls[r] == (int)kReplayGlobal) {
            const int j = (r + D + 1 - rot) % (D + 1);
            const uint32_t gs = rp.gs[j * rp.stride + rq];
            const int id = gs <= ta.mask ? slot_vertex(ta, gs) : -1;
            if (id >= 0 && (unsigned)id < M) {
                float row[kMaxVD];
                load_row<T, LVD>(vals, id, row);
#pragma unroll
                for (int c = 0; c < LVD; c++) acc[c] = acc[c] + s.b[r] * (row[c] * un[c]);
            }
        }
    }
    write_pixel<LVD>(tv.dst, p, vd, acc);
}

// Slice, direct form (D > 6 or vd > 4): per-pixel global lookups, the pixels
// in XCD bands.
template <int D, bool PACKED, bool BILAT, typename T>
__global__ __launch_bounds__(kBlock) void k_slice_direct(PixSrc src, int n, int vd, int vs, LatConst lc, KeyPlan kp,
                                                         TableArgs ta, const T *__restrict__ vals,
                                                         const unsigned *__restrict__ sbits, OutDst dst,
                                                         Counters *ctr, int count) {
    using K = KeyOps<D, PACKED>;
    constexpr int W = K::W;
    if (*(volatile const unsigned *)&ctr->flags & kFlagFull) return;   // redone with a larger table
    const int p = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    uint32_t probes = 0;
    unsigned flags = 0;
    if (p < n) {
        float pos[D];
        load_position<D, BILAT>(src, p, pos);
        Simplex<D> s;
        simplex<D>(pos, lc.sf, s);
        float acc[kMaxVD];
#pragma unroll
        for (int c = 0; c < kMaxVD; c++) acc[c] = 0.0f;
        if (s.bad) {
            flags |= kFlagBad;
        } else {
            float un[kMaxVD];
            slice_unscale<T, kMaxVD>(sbits, un);
            uint64_t pk[D + 1];
            bool all = false;
            if constexpr (PACKED) all = packed_vertex_keys<D>(s, kp, 0ull, pk);
#pragma unroll
            for (int r = 0; r <= D; r++) {
                uint64_t k[W];
                if (PACKED && all) k[0] = pk[r];
                else if (!K::make(s, r, kp, 0ull, k)) continue;   // outside the plan: absent
                const int id = table_vertex<D, PACKED, false>(ta, kp, k, probes);
                if (id < 0) continue;
                float v[kMaxVD];
                load_row_rt<T>(vals, id, vs, v);
                if (!std::is_same<T, float>::value) {
#pragma unroll
                    for (int c = 0; c < kMaxVD; c++) v[c] *= un[c];
                }
#pragma unroll
                for (int c = 0; c < kMaxVD; c++) acc[c] = acc[c] + s.b[r] * v[c];
            }
            write_pixel<kMaxVD>(dst, p, vd, acc);
        }
    }
    report_flags(ctr, flags);
    if (count) report_counts(ctr, probes, 0, 0);
}

template <int D, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_unpack_keys(KeyPlan kp, const uint64_t *__restrict__ vkeys, int M,
                                                        int16_t *__restrict__ keys) {
    using K = KeyOps<D, PACKED>;
    constexpr int W = K::W;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M) return;
    uint64_t k[W];
#pragma unroll
    for (int w = 0; w < W; w++) k[w] = vkeys[(size_t)i * W + w];
#pragma unroll
    for (int q = 0; q < D; q++) keys[(size_t)i * D + q] = (int16_t)K::coord(k, q, kp);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_store_to_f32(const T *__restrict__ vals, const unsigned *__restrict__ sbits,
                                                         int M, int vd, int vs, float *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M) return;
    float un[kMaxVD];
    slice_unscale<T, kMaxVD>(sbits, un);
#pragma unroll
    for (int c = 0; c < kMaxVD; c++)
        if (c < vd) out[(size_t)i * vd + c] = to_f32<T>(vals[(size_t)i * vs + c]) * un[c];
}

} // namespace gt
