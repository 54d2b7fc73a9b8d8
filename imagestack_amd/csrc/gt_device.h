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

namespace gt {

constexpr int kMaxD = 16;
constexpr int kMaxVD = 16;
constexpr uint64_t kEmpty = ~0ull;     // an unused table word (no valid key packs to it)
constexpr int kWave = 64;

// Number of 64-bit words in a packed key of D int16 coordinates.
template <int D> struct KW { static constexpr int value = (D + 3) / 4; };

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

    // Walk back onto the hyperplane (:326-348).
    if (sum > 0) {
#pragma unroll
        for (int i = 0; i <= D; i++) {
            const bool hi = s.rk[i] >= D + 1 - sum;
            s.g[i] -= hi ? (D + 1) : 0;
            s.rk[i] += hi ? sum - (D + 1) : sum;
        }
    } else if (sum < 0) {
#pragma unroll
        for (int i = 0; i <= D; i++) {
            const bool lo = s.rk[i] < -sum;
            s.g[i] += lo ? (D + 1) : 0;
            s.rk[i] += lo ? (D + 1) + sum : sum;
        }
    }

    // Barycentric weights (:350-356).  The reference accumulates
    // b[D-rk_i] += delta_i and b[D+1-rk_i] -= delta_i in index order.  Ranks
    // are a permutation, so each b[m] receives exactly one + and one - term
    // and, IEEE addition being commutative, b[m] = delta(rank D-m) -
    // delta(rank D+1-m) bit for bit, whichever term came first.  delta by
    // rank is selected without dynamic register indexing.
    float dr[D + 1];
#pragma unroll
    for (int k = 0; k <= D; k++) dr[k] = 0.0f;
#pragma unroll
    for (int i = 0; i <= D; i++) {
        const float di = (e[i] - (float)s.g[i]) * scale;
#pragma unroll
        for (int k = 0; k <= D; k++) dr[k] = (s.rk[i] == k) ? di : dr[k];
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

// Packed key: coordinate i biased to uint16 (k + 32768) in word i/4, bits
// 48-16*(i%4).  Coordinate 0 is the most significant, so comparing words in
// order as unsigned integers is the lexicographic order of the signed keys.
// Keys are kept within +-(32767 - 2(D+1)), so no word equals kEmpty.
template <int D>
__device__ __forceinline__ void pack_vertex(const Simplex<D> &s, int r, uint64_t (&k)[KW<D>::value]) {
#pragma unroll
    for (int w = 0; w < KW<D>::value; w++) k[w] = 0;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const uint64_t u = (uint64_t)(uint16_t)(vertex_coord(s, r, i) + 32768);
        k[i / 4] |= u << (48 - 16 * (i % 4));
    }
}

template <int D>
__device__ __forceinline__ int key_coord(const uint64_t (&k)[KW<D>::value], int i) {
    return (int)((k[i / 4] >> (48 - 16 * (i % 4))) & 0xffff) - 32768;
}

template <int D>
__device__ __forceinline__ void set_key_coord(uint64_t (&k)[KW<D>::value], int i, int c) {
    const int sh = 48 - 16 * (i % 4);
    k[i / 4] = (k[i / 4] & ~(0xffffull << sh)) | ((uint64_t)(uint16_t)(c + 32768) << sh);
}

template <int D>
__device__ __forceinline__ uint64_t hash_key(const uint64_t (&k)[KW<D>::value]) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int w = 0; w < KW<D>::value; w++) {
        h ^= k[w];
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 33;
    }
    h *= 0xc4ceb9fe1a85ec53ull;
    return h ^ (h >> 29);
}

// Insert-or-find in the global open-addressing table (linear probing, slot
// s owns words tkeys[s*KW .. s*KW+KW-1]).  Each word is claimed by its own
// 64-bit CAS from kEmpty; a thread whose word is already there moves on to
// the next word, and a mismatch on any word sends it to the next slot.
// Words are written once, so every thread with a given key agrees on the
// slot that holds it, without locks or spinning (no intra-wave deadlock).
// The thread whose CAS completes the last word of a slot is its creator.
// Returns the slot, or -1 if max_probe slots were visited.
template <int D>
__device__ __forceinline__ int table_insert(uint64_t *__restrict__ tkeys, uint32_t mask,
                                            const uint64_t (&k)[KW<D>::value], uint64_t h,
                                            bool &created, uint32_t &probes) {
    constexpr int W = KW<D>::value;
    uint32_t s = (uint32_t)h & mask;
    created = false;
    for (uint32_t it = 0; it <= mask; it++) {
        probes++;
        bool match = true;
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (match) {
                uint64_t *a = tkeys + (size_t)s * W + w;
                uint64_t cur = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == kEmpty) {
                    uint64_t expect = kEmpty;
                    const bool ok = __hip_atomic_compare_exchange_strong(
                        a, &expect, k[w], __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    cur = ok ? k[w] : expect;
                    if (ok && w == W - 1) created = true;
                }
                match = (cur == k[w]);
            }
        }
        if (match) return (int)s;
        s = (s + 1) & mask;
    }
    return -1;
}

// Read-only lookup after the table is complete (a later kernel): plain
// loads, returns the slot or -1.
template <int D>
__device__ __forceinline__ int table_find(const uint64_t *__restrict__ tkeys, uint32_t mask,
                                          const uint64_t (&k)[KW<D>::value], uint64_t h,
                                          uint32_t &probes) {
    constexpr int W = KW<D>::value;
    uint32_t s = (uint32_t)h & mask;
    for (uint32_t it = 0; it <= mask; it++) {
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

// Lattice value storage.
template <typename T> struct Store;
template <> struct Store<float> {
    static __device__ __forceinline__ float load(const float *p) { return *p; }
    static __device__ __forceinline__ void store(float *p, float v) { *p = v; }
};
template <> struct Store<__hip_bfloat16> {
    static __device__ __forceinline__ float load(const __hip_bfloat16 *p) { return __bfloat162float(*p); }
    static __device__ __forceinline__ void store(__hip_bfloat16 *p, float v) { *p = __float2bfloat16(v); }
};
template <> struct Store<__half> {
    static __device__ __forceinline__ float load(const __half *p) { return __half2float(*p); }
    static __device__ __forceinline__ void store(__half *p, float v) { *p = __float2half(v); }
};

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

} // namespace gt
