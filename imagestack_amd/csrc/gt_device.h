// Device-side building blocks of the lattice kernels (gfx950, wave64).
//
// Compile with -ffp-contract=off: the elevation and barycentric arithmetic
// must round exactly like the reference's scalar FP32 code so that lattice
// keys and weights are bit-identical (src/Permutohedral.h:293-356).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "gt_internal.h"

namespace gt {

constexpr int kMaxD = 16;
constexpr int kMaxVD = 16;
constexpr uint64_t kEmpty = ~0ull;     // an unused table word (no valid key word equals it)
constexpr int kWave = 64;
constexpr uint32_t kMaxProbes = 256;   // an insert that probes this far reports a full table
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));   // a 16-byte vector access
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));   // an 8-byte vector access

// Enclosing simplex of one point: fixed-up remainder-0 point g, ranks rk and
// barycentric weights b[0..D].  `bad` flags NaN/Inf or keys beyond int16.
template <int D>
struct Simplex {
    int g[D + 1];
    int rk[D + 1];
    float b[D + 1];
    bool bad;
};

// src/Permutohedral.h:293-356, restated for registers.  Every array index is
// a compile-time constant after unrolling, so nothing spills to scratch.
template <int D>
__device__ __forceinline__ void simplex(const float (&p)[D], const float *__restrict__ sf,
                                        Simplex<D> &s) {
    float e[D + 1];
    // Elevation onto the hyperplane sum(x) = 0 (:293-298).
    e[D] = ((float)(-D) * p[D - 1]) * sf[D - 1];
#pragma unroll
    for (int i = D - 1; i > 0; i--)
        e[i] = (e[i + 1] - ((float)i * p[i - 1]) * sf[i - 1]) + ((float)(i + 2) * p[i]) * sf[i];
    e[0] = e[1] + (2.0f * p[0]) * sf[0];

    // Nearest remainder-0 point, ties rounding down (:305-317).
    const float scale = 1.0f / (float)(D + 1);
    int sum = 0;
    bool bad = false;
#pragma unroll
    for (int i = 0; i <= D; i++) {
        const float v = e[i] * scale;
        const float up = ceilf(v) * (float)(D + 1);
        const float dn = floorf(v) * (float)(D + 1);
        const float gf = (up - e[i] < e[i] - dn) ? up : dn;
        bad |= !(fabsf(gf) <= 32767.0f - (float)(2 * (D + 1)));
        s.g[i] = (int)gf;
        sum += s.g[i];
    }
    sum /= (D + 1);

    // Ranks of the differential; ties go to the later index (:321-324).
    float diff[D + 1];
#pragma unroll
    for (int i = 0; i <= D; i++) { diff[i] = e[i] - (float)s.g[i]; s.rk[i] = 0; }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = i + 1; j <= D; j++) {
            const bool lt = diff[i] < diff[j];
            s.rk[i] += lt ? 1 : 0;
            s.rk[j] += lt ? 0 : 1;
        }

    // Walk back onto the hyperplane (:326-348).  For sum > 0 the reference
    // moves every i with rk_i >= D+1-sum down by D+1 and takes D+1 from its
    // shifted rank; for sum < 0 every i with rk_i < -sum up.  With t = rk_i +
    // sum those are exactly t >= D+1 and t < 0 (neither happens for the other
    // sign, nor for sum = 0), which needs no branch on the sign of sum.
#pragma unroll
    for (int i = 0; i <= D; i++) {
        const int t = s.rk[i] + sum;
        const int w = (t < 0) ? (D + 1) : ((t >= D + 1) ? -(D + 1) : 0);
        s.g[i] += w;
        s.rk[i] = t + w;
    }

    // Barycentric weights (:350-356).  The reference accumulates
    // b[D-rk_i] += delta_i and b[D+1-rk_i] -= delta_i in index order.  Ranks
    // are a permutation, so each b[m] receives exactly one + and one - term
    // and, IEEE addition being commutative, b[m] = delta(rank D-m) -
    // delta(rank D+1-m) bit for bit, whichever term came first.  delta by
    // rank is the deltas in descending order: the ranks count the larger
    // differentials (ties to the later index), and the walk back moves the
    // wrapped coordinates' deltas up by exactly 1 -- above every other, the
    // differentials spanning at most D+1 -- together with their ranks to the
    // front.  Equal deltas are equal values, so an odd-even transposition
    // network of max/min gives the same dr as selecting by rank, in
    // D(D+1)/2 comparators.  (Only where the greedy rounding picked the
    // farther multiple by a rounding error can a wrapped delta fall an ulp
    // below another; the two weights it feeds then differ by that ulp.)
    float dr[D + 1];
#pragma unroll
    for (int i = 0; i <= D; i++) dr[i] = (e[i] - (float)s.g[i]) * scale;
#pragma unroll
    for (int round = 0; round <= D; round++) {
#pragma unroll
        for (int i = round & 1; i + 1 <= D; i += 2) {
            const float a = dr[i], b = dr[i + 1];
            dr[i] = fmaxf(a, b);
            dr[i + 1] = fminf(a, b);
        }
    }
    s.b[0] = dr[D] + (1.0f - dr[0]);
#pragma unroll
    for (int m = 1; m <= D; m++) s.b[m] = dr[D - m] - dr[D + 1 - m];
    s.bad = bad;
}

// Coordinate i of vertex r (:365-369): g_i + canonical[r][rk_i] with
// canonical[r][j] = r for j <= D-r and r-(D+1) otherwise (:247-254).
template <int D>
__device__ __forceinline__ int vertex_coord(const Simplex<D> &s, int r, int i) {
    return s.g[i] + ((s.rk[i] <= D - r) ? r : r - (D + 1));
}

// ---------------------------------------------------------------------------
// Key formats.  Both keep coordinate 0 most significant, so comparing the
// words from the highest index down as unsigned integers is the
// lexicographic order of the signed keys, and neither lets a word equal
// kEmpty.
//
// Packed: one word; coordinate i stored as (c_i - lo_i) in `bits_i` bits at
// `shift_i` (host-computed KeyPlan, bit 63 always clear).  Used whenever the
// plan fits 63 bits.
//
// Wide: ceil(D/4) words; coordinate i biased to uint16 (c + 32768) at bits
// 48-16*(i%4) of word i/4.  A key coordinate stays within +-(32767 - 2(D+1)),
// so the first field of a word is never 0xffff.
template <int D, bool PACKED> struct KeyOps;

// `tag` holds fields above the coordinates: the image index of a batched
// lattice (packed format only; 0 otherwise).
template <int D>
struct KeyOps<D, true> {
    static constexpr int W = 1;
    __device__ static __forceinline__ bool from_coords(const int (&c)[D], const KeyPlan &kp, uint64_t tag,
                                                       uint64_t (&k)[W]) {
        uint64_t v = tag;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < D; i++) {
            const uint32_t u = (uint32_t)(c[i] - kp.lo[i]);
            ok &= u <= kp.span[i];
            v |= (uint64_t)(u & kp.span[i]) << kp.shift[i];
        }
        k[0] = v;
        return ok;
    }
    __device__ static __forceinline__ bool make(const Simplex<D> &s, int r, const KeyPlan &kp, uint64_t tag,
                                                uint64_t (&k)[W]) {
        int c[D];
#pragma unroll
        for (int i = 0; i < D; i++) c[i] = vertex_coord(s, r, i);
        return from_coords(c, kp, tag, k);
    }
    __device__ static __forceinline__ int coord(const uint64_t (&k)[W], int i, const KeyPlan &kp) {
        return (int)((k[0] >> kp.shift[i]) & kp.span[i]) + kp.lo[i];
    }
    __device__ static __forceinline__ uint64_t tag_of(const uint64_t (&k)[W], const KeyPlan &kp) {
        return kp.coord_bits >= 64 ? 0ull : (k[0] >> kp.coord_bits) << kp.coord_bits;
    }
};

template <int D>
struct KeyOps<D, false> {
    static constexpr int W = (D + 3) / 4;
    __device__ static __forceinline__ bool from_coords(const int (&c)[D], const KeyPlan &kp, uint64_t,
                                                       uint64_t (&k)[W]) {
        bool ok = true;
#pragma unroll
        for (int w = 0; w < W; w++) k[w] = 0;
#pragma unroll
        for (int i = 0; i < D; i++) {
            ok &= (c[i] >= -32767 + 2 * (D + 1)) && (c[i] <= 32767 - 2 * (D + 1));
            k[i / 4] |= (uint64_t)(uint16_t)(c[i] + 32768) << (48 - 16 * (i % 4));
        }
        return ok;
    }
    __device__ static __forceinline__ bool make(const Simplex<D> &s, int r, const KeyPlan &kp, uint64_t tag,
                                                uint64_t (&k)[W]) {
        int c[D];
#pragma unroll
        for (int i = 0; i < D; i++) c[i] = vertex_coord(s, r, i);
        return from_coords(c, kp, tag, k);
    }
    __device__ static __forceinline__ int coord(const uint64_t (&k)[W], int i, const KeyPlan &) {
        return (int)((k[i / 4] >> (48 - 16 * (i % 4))) & 0xffff) - 32768;
    }
    __device__ static __forceinline__ uint64_t tag_of(const uint64_t (&)[W], const KeyPlan &) { return 0ull; }
};

// Packed keys of all D+1 vertices of a simplex at once (src/Permutohedral.h:
// 365-369 for every r).  Coordinate i of vertex r is g_i + r - (D+1)[rk_i >=
// D+1-r]: vertex 0 is g, and each next vertex adds 1 to every coordinate and
// takes D+1 from the one coordinate of rank D+1-r (none when that is the
// implicit coordinate D).  In the packed word that is key(r) = key(r-1) +
// ones - drop_r, exact while every field stays in its range: coordinate i
// spans [g_i - rk_i, g_i - rk_i + D] over the simplex, so one check per
// coordinate covers all vertices.  Returns false (keys unusable) when a
// coordinate leaves the plan.
template <int D>
__device__ __forceinline__ bool packed_vertex_keys(const Simplex<D> &s, const KeyPlan &kp, uint64_t tag,
                                                   uint64_t (&key)[D + 1]) {
    uint64_t k = tag, ones = 0;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const int lo_f = s.g[i] - s.rk[i] - kp.lo[i];   // the field's smallest value over the vertices
        ok &= lo_f >= 0 && lo_f + D <= (int)kp.span[i];
        k += (uint64_t)(uint32_t)(s.g[i] - kp.lo[i]) << kp.shift[i];
        ones += 1ull << kp.shift[i];
    }
    key[0] = k;
#pragma unroll
    for (int r = 1; r <= D; r++) {
        uint64_t drop = 0;
#pragma unroll
        for (int i = 0; i < D; i++) drop = (s.rk[i] == D + 1 - r) ? ((uint64_t)(D + 1) << kp.shift[i]) : drop;
        k = k + ones - drop;
        key[r] = k;
    }
    return ok;
}

// Key hash: the words mixed one by one and summed, then a final mix.  A
// sum, so that a key that differs from a known one in a single word (a
// lattice neighbour: the same constant added to every field, then one field
// moved) hashes with one word mix instead of W (k_nbr_hash_vertex).
__device__ __forceinline__ uint64_t hash_word(uint64_t x, int w) {
    x ^= 0x9E3779B97F4A7C15ull * (uint64_t)(2 * w + 1);
    x *= 0xff51afd7ed558ccdull;
    return x ^ (x >> 33);
}
__device__ __forceinline__ uint64_t hash_final(uint64_t h) {
    h *= 0xc4ceb9fe1a85ec53ull;
    return h ^ (h >> 29);
}
template <int W>
__device__ __forceinline__ uint64_t hash_key(const uint64_t (&k)[W]) {
    uint64_t h = 0;
#pragma unroll
    for (int w = 0; w < W; w++) h += hash_word(k[w], w);
    return hash_final(h);
}

// Insert-or-find in the global open-addressing table (linear probing; slot s
// owns words tkeys[s*W .. s*W+W-1]), words written once.  Returns the slot,
// -1 after kMaxProbes slots (table full), or -2 when a wide key's words
// never appeared (kFlagWait).
//
// One-word keys: a workgroup-scope (cache-friendly) read, then a CAS from
// kEmpty only where the word looks empty.
//
// Wide keys: the slot belongs to whoever claims word 0 by CAS; the claimer
// then stores words 1..W-1 (device-coherent stores, no atomics).  A thread
// that finds word 0 equal to its own reads the other words, waiting while
// one still reads kEmpty -- its claimer has already issued the store: in a
// wave the claiming lanes store before any lane reaches the wait, and a
// claimer elsewhere stores right after its CAS returned -- then compares;
// any mismatch moves on to the next slot.  A new vertex costs one CAS and
// W-1 stores instead of W dependent read-and-CAS pairs (d = 16: 17 new
// vertices a pixel).  The wait is bounded: past kWaitSpins reads it reports
// -2 rather than spin forever.
constexpr uint32_t kWaitSpins = 1u << 20;
template <int W>
__device__ __forceinline__ int table_insert(uint64_t *__restrict__ tkeys, uint32_t mask, const uint64_t (&k)[W],
                                            uint32_t s, uint32_t &probes) {
    if constexpr (W > 1) {
        for (uint32_t it = 0; it < kMaxProbes; it++) {
            probes++;
            uint64_t *a = tkeys + (size_t)s * W;
            uint64_t cur = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == kEmpty) {
                uint64_t expect = kEmpty;
                __hip_atomic_compare_exchange_strong(a, &expect, k[0], __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                if (expect == kEmpty) {   // claimed: the other words are this key's
#pragma unroll
                    for (int w = 1; w < W; w++)
                        __hip_atomic_store(a + w, k[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return (int)s;
                }
                cur = expect;
            }
            if (cur == k[0]) {
                bool match = true;
#pragma unroll
                for (int w = 1; w < W; w++) {
                    uint64_t v = __hip_atomic_load(a + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    for (uint32_t n = 0; v == kEmpty; n++) {
                        if (n >= kWaitSpins) return -2;
                        v = __hip_atomic_load(a + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    match = match && v == k[w];
                }
                if (match) return (int)s;
            }
            s = (s + 1) & mask;
        }
        return -1;
    }
    for (uint32_t it = 0; it < kMaxProbes; it++) {
        probes++;
        bool match = true;
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (match) {
                uint64_t *a = tkeys + (size_t)s * W + w;
                // A workgroup-scope (non-coherent, cache-friendly) read first:
                // words are write-once, so a value other than kEmpty is final;
                // only an apparently empty word needs the device-scope CAS.
                uint64_t cur = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cur == kEmpty) {
                    uint64_t expect = kEmpty;
                    __hip_atomic_compare_exchange_strong(a, &expect, k[w], __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                    // expect now holds the previous word: kEmpty if our CAS wrote it.
                    cur = (expect == kEmpty) ? k[w] : expect;
                }
                match = (cur == k[w]);
            }
        }
        if (match) return (int)s;
        s = (s + 1) & mask;
    }
    return -1;
}

// Flag of a failed table_insert.
__device__ __forceinline__ unsigned insert_fail_flag(int slot) { return slot == -2 ? kFlagWait : kFlagFull; }

// Tag of a key with 64-bit hash h (TableArgs::tags): never 0 (empty).
__device__ __forceinline__ uint8_t key_tag8(uint64_t h) { return (uint8_t)(0x80u | (uint32_t)(h >> 57)); }

// The tag of key k at its slot (tables with tags: wide keys), written by
// every insert that found or claimed the slot (the same byte each time), so
// that the count and the finalize can read a slot's occupancy from one byte.
template <int W>
__device__ __forceinline__ void write_tag(const TableArgs &ta, int slot, const uint64_t (&k)[W]) {
    if constexpr (W > 1) {
        if (ta.tags && slot >= 0) ta.tags[slot] = key_tag8(hash_key<W>(k));
    }
}

// Lookup through the tags (TableArgs::tags) of a complete table, from home
// slot s: eight tags at a time (two aligned 8-byte loads), the empty ones and
// the ones equal to `want` found by byte-parallel masks, key words compared
// only at candidates before the first empty slot, in probe order.  The
// zero-byte test may also flag a byte just above a zero one; a flagged byte
// is a candidate only, checked against the key words, so the result is
// exact.  Returns the slot (w0: its first key word) or -1.
template <int W>
__device__ __forceinline__ int tag_find(const TableArgs &ta, const uint64_t (&k)[W], uint64_t kmask, uint32_t s,
                                        uint8_t want, uint64_t &w0) {
    constexpr uint64_t kOnes = 0x0101010101010101ull, kTops = 0x8080808080808080ull;
    const uint64_t wv = kOnes * want;
    for (uint32_t n = 0; n < kMaxProbes; n += 8) {
        const uint32_t a = s & ~7u;
        const uint64_t lo = *reinterpret_cast<const uint64_t *>(ta.tags + a);
        const uint64_t hi = *reinterpret_cast<const uint64_t *>(ta.tags + ((a + 8) & ta.mask));
        const int sh = 8 * (int)(s & 7u);
        const uint64_t t = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;   // tags of slots s .. s+7
        const uint64_t empty = ~t & kTops;
        const uint64_t x = t ^ wv;
        uint64_t cand = (x - kOnes) & ~x & kTops;
        if (empty) cand &= (empty & (0ull - empty)) - 1;   // before the first empty slot
        while (cand) {
            const uint32_t slot = (s + (uint32_t)(__builtin_ctzll(cand) >> 3)) & ta.mask;
            const uint64_t *kw = ta.tkeys + (size_t)slot * W;
            const uint64_t v0 = kw[0];
            bool match = (v0 & kmask) == k[0];
#pragma unroll
            for (int w = 1; w < W; w++) match = match && kw[w] == k[w];
            if (match) {
                w0 = v0;
                return (int)slot;
            }
            cand &= cand - 1;
        }
        if (empty) return -1;
        s = (s + 8) & ta.mask;
    }
    return -1;
}

// Read-only lookup in a complete table (a later kernel): plain loads.
template <int W>
__device__ __forceinline__ int table_find(const uint64_t *__restrict__ tkeys, uint32_t mask, const uint64_t (&k)[W],
                                          uint32_t s, uint32_t &probes) {
    for (uint32_t it = 0; it < kMaxProbes; it++) {
        probes++;
        const uint64_t *a = tkeys + (size_t)s * W;
        const uint64_t w0 = a[0];
        if (w0 == kEmpty) return -1;
        bool match = (w0 == k[0]);
#pragma unroll
        for (int w = 1; w < W; w++) match = match && (a[w] == k[w]);
        if (match) return (int)s;
        s = (s + 1) & mask;
    }
    return -1;
}

// Home slot in a partitioned table of a key with hash h whose fields sum to
// `fields` (part_base offsets that sum to c_0 + ... + c_{D-1} + margin) and
// whose image tag is `tag`.  BATCH: the caller may be working on a batch
// (per-image slices, TableArgs::part_img); single-image kernels compile
// without that path.
template <bool BATCH = true>
__device__ __forceinline__ uint32_t slot_in_part(const TableArgs &ta, uint32_t h, int fields, uint32_t tag) {
    uint32_t part;
    if (BATCH && ta.part_img) {
        const BatchImage &bi = ta.part_img[tag];
        const uint32_t u = (uint32_t)min(max(fields + ta.part_base, 0), bi.part_range);
        const uint64_t pos24 = bi.part_start24 + (((uint64_t)u * bi.part_mulw) >> 24);
        part = (uint32_t)((pos24 << ta.part_lg) >> 24);
    } else {
        const uint32_t u = (uint32_t)min(max(fields + ta.part_base, 0), ta.part_range);
        part = (uint32_t)(((uint64_t)u * ta.part_mul) >> 24);
    }
    return (part << ta.part_shift) | (h & (ta.mask >> ta.part_lg));
}

// Image tag of a one-word key (a batch's keys carry it above the fields).
__device__ __forceinline__ uint32_t key_tag(const TableArgs &ta, const KeyPlan &kp, uint64_t w0) {
    return ta.part_img ? (uint32_t)(w0 >> kp.coord_bits) : 0u;
}

// Home slot of key k (where its linear probe starts): the hash, or in a
// partitioned table (TableArgs::part_mul) the hash inside the key's part.
template <int D, bool PACKED, bool BATCH = true>
__device__ __forceinline__ uint32_t home_slot(const TableArgs &ta, const KeyPlan &kp,
                                              const uint64_t (&k)[KeyOps<D, PACKED>::W]) {
    constexpr int W = KeyOps<D, PACKED>::W;
    const uint32_t h = (uint32_t)hash_key<W>(k);
    if (ta.part_mul) {
        int sum = 0;
#pragma unroll
        for (int i = 0; i < D; i++)
            sum += PACKED ? (int)((k[0] >> kp.shift[i]) & kp.span[i])
                          : (int)((k[i / 4] >> (48 - 16 * (i % 4))) & 0xffff);
        return slot_in_part<BATCH>(ta, h, sum, (PACKED && BATCH) ? key_tag(ta, kp, k[0]) : 0u);
    }
    return h & ta.mask;
}

// Vertex id of key k in a finalized table (after k_finalize), or -1: read
// from the key's own word when the ids were folded in (TableArgs::idshift),
// else through slot_id.
template <int D, bool PACKED, bool BATCH>
__device__ __forceinline__ int table_vertex(const TableArgs &ta, const KeyPlan &kp,
                                            const uint64_t (&k)[KeyOps<D, PACKED>::W], uint32_t &probes) {
    constexpr int W = KeyOps<D, PACKED>::W;
    uint32_t s = home_slot<D, PACKED, BATCH>(ta, kp, k);
    if (W == 1 && ta.idshift > 0) {
        for (uint32_t it = 0; it < kMaxProbes; it++) {
            probes++;
            const uint64_t w0 = ta.tkeys[s];
            if (w0 == kEmpty) return -1;
            if ((w0 & ta.keymask) == k[0]) return (int)(w0 >> ta.idshift);
            s = (s + 1) & ta.mask;
        }
        return -1;
    }
    const int slot = table_find<W>(ta.tkeys, ta.mask, k, s, probes);
    return slot >= 0 ? ta.slot_id[slot] : -1;
}

// Workgroup-local key table in LDS: the distinct vertices of one pixel tile.
// Same word-claim protocol as the global table.  Returns the LDS slot, or -1
// after MAXP probes (a tile with more distinct vertices than the table holds
// then sends the rest straight to the global table).
//
// The slot hash of a one-word key is two 32-bit multiplies (a tile's keys
// differ in a few low fields; the fold and the odd multipliers spread them);
// wider keys use hash_key.
template <int W, int LT>
__device__ __forceinline__ uint32_t lds_home(const uint64_t (&k)[W]) {
    if (W == 1) {
        const uint32_t x = (uint32_t)k[0] ^ ((uint32_t)(k[0] >> 32) * 0x85EBCA6Bu);
        return (x * 0x9E3779B1u) >> (32 - __builtin_ctz((unsigned)LT));
    }
    return (uint32_t)(hash_key<W>(k) >> 40) & (LT - 1);
}

//
// The probe loop stays a loop: unrolled (MAXP times, for each of a pixel's
// D+1 vertices) every level of it holds its own saved exec mask, and the
// SGPRs those take spill the kernel's arguments into VGPR lanes.
template <int W, int LT, int MAXP>
__device__ __forceinline__ int lds_insert(uint64_t *lk, const uint64_t (&k)[W]) {
    uint32_t ls = lds_home<W, LT>(k);
#pragma unroll 1
    for (int it = 0; it < MAXP; it++) {
        bool match = true;
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (match) {
                uint64_t expect = kEmpty;
                __hip_atomic_compare_exchange_strong(lk + ls * W + w, &expect, k[w], __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                match = (expect == kEmpty) || (expect == k[w]);
            }
        }
        if (match) return (int)ls;
        ls = (ls + 1) & (LT - 1);
    }
    return -1;
}

// x * s rounded to the nearest integer (ties to even), as a two's-complement
// 64-bit word; |x * s| < 2^62.  y = rint(x * s) is an integer-valued float
// with at most 24 significant bits, so hi = floor(y / 2^32) and the low word
// lo = y - hi * 2^32 in [0, 2^32) are exact for y >= 0 and for y <= -2^31,
// giving the same value as a float -> int64 conversion without its sign
// handling; for -2^31 < y < 0, lo = 2^32 + y rounds to FP32 and the word lies
// within 2^8 units of y (deterministic all the same).
__device__ __forceinline__ unsigned long long to_fixed(float x, float s) {
    const float y = rintf(x * s);
    const float hi = floorf(y * 2.3283064365386963e-10f);   // 2^-32
    const float lo = fmaf(hi, -4294967296.0f, y);
    return ((unsigned long long)(uint32_t)(int)hi << 32) | (uint32_t)lo;
}

// Exponent e of a narrow storage's scale 2^e for values of largest finite
// magnitude m (float bits; 0: all zero): e = 14 - ceil(log2(m)), exact
// through frexpf, clamped to [-100, 100].  Scaled values stay below 2^14.
__device__ __forceinline__ int store_exp(unsigned mbits) {
    if (mbits == 0) return 0;
    int k = 0;
    const float f = frexpf(__uint_as_float(mbits), &k);   // m = f 2^k, f in [0.5, 1)
    const int lg = (f == 0.5f) ? k - 1 : k;
    return min(100, max(-100, 14 - lg));
}

// Running maximum of the finite magnitudes of o[0..NV) into m.
template <int NV>
__device__ __forceinline__ void take_max(float (&m)[kMaxVD], const float (&o)[kMaxVD]) {
#pragma unroll
    for (int c = 0; c < NV; c++) {
        const float a = fabsf(o[c]);
        m[c] = (a <= 3.402823466e38f) ? fmaxf(m[c], a) : m[c];
    }
}

// Workgroup maximum of m[c] per channel c < nc into track[c] (float bits:
// non-negative floats order as their bits).  Waves meet in LDS, then one
// thread per channel raises the global word, and only when the workgroup's
// maximum exceeds what the word already holds: a pass over millions of
// vertices launches ~10^5 workgroups, and same-address global atomics from
// all of them serialise (the maximum only grows, so a stale read can cost an
// atomic but never skip a needed one).  Every thread of the workgroup must
// call it.
template <int NV>
__device__ __forceinline__ void publish_max(unsigned *track, const float (&m)[kMaxVD], int nc) {
    __shared__ unsigned wg_max[NV];
    if (threadIdx.x < NV) wg_max[threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NV; c++) {
        if (c >= nc) continue;
        float v = m[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
        if (__lane_id() == 0 && v > 0.0f) atomicMax(&wg_max[c], __float_as_uint(v));
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)nc && threadIdx.x < NV) {
        const unsigned b = wg_max[threadIdx.x];
        if (b > __hip_atomic_load(track + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(track + threadIdx.x, b);
    }
}

// The fixed-point word of a product already in fixed-point units, y = (w v) s
// with |y| < 2^51, computed as w * (v s) (s a power of two, so the product
// rounds once, as (w v) s would): rint(y) as a two's-complement word.  In
// FP64, 1.5 * 2^52 + y lies in [2^52, 2^53), where the spacing is 1, so the
// sum rounds y to the nearest integer and its bit pattern is that of
// 1.5 * 2^52 plus rint(y): one conversion, one add, one 64-bit subtract.
__device__ __forceinline__ unsigned long long fixed_word(float y) {
    const double t = (double)y + 6755399441055744.0;   // 1.5 * 2^52
    return (unsigned long long)__double_as_longlong(t) - 0x4338000000000000ull;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Storage conversions of lattice values.
template <typename T> __device__ __forceinline__ float to_f32(T v);
template <> __device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f32<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <> __device__ __forceinline__ float to_f32<__half>(__half v) { return __half2float(v); }
template <typename T> __device__ __forceinline__ T from_f32(float v);
template <> __device__ __forceinline__ float from_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ __hip_bfloat16 from_f32<__hip_bfloat16>(float v) { return __float2bfloat16(v); }
template <> __device__ __forceinline__ __half from_f32<__half>(float v) { return __float2half(v); }

// Vertex value rows: vertex-major, VS (a power of two >= vd) values each.
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

template <typename T>
__device__ __forceinline__ void store_row_rt(T *base, int64_t i, int vs, const float (&o)[kMaxVD]) {
    switch (vs) {
    case 1: store_row<T, 1>(base, i, o); break;
    case 2: store_row<T, 2>(base, i, o); break;
    case 4: store_row<T, 4>(base, i, o); break;
    case 8: store_row<T, 8>(base, i, o); break;
    default: store_row<T, 16>(base, i, o); break;
    }
}

} // namespace gt
