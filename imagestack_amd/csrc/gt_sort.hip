// LSD radix sort of (multi-word key, uint32 payload) pairs for the lattice
// vertex keys, hand-written for gfx950 (wave64): one pass per 8-bit digit
// with decoupled look-back (a "one-sweep" radix sort).
//
//   k_sort_init   : new epoch, tile counters and digit counts to zero.
//   k_sort_counts : the digit counts of every pass in one read of the keys
//                   (an LSD pass permutes the keys but not their digits, so
//                   every pass's global digit offsets are known up front).
//   k_sort_pass   : per 8-bit digit, one workgroup per tile (tile = workgroup
//                   index: workgroups are dispatched in index order, so every
//                   tile a look-back waits for is running or done; a shared
//                   atomic tile counter instead costs ~50 ns per workgroup,
//                   serialised).  Each ranks its tile stably (each wave
//                   owns a run of consecutive keys and ranks 64 at a time
//                   with eight ballots), publishes its per-digit counts,
//                   looks back over earlier tiles for their running totals,
//                   orders the tile by digit in LDS and writes it out in
//                   digit runs.
//
// The item count lives on the device (Counters::M): grids are sized for the
// host's upper bound and surplus workgroups exit at once.  The look-back
// status words carry a tag (epoch * 32 + pass), so they never need clearing;
// a look-back that waits too long (it cannot, short of a bug) raises
// kFlagSort instead of spinning forever.
#include "gt_device.h"
#include "gt_internal.h"

#include <utility>

namespace gt {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kMaxPasses = 8 * 4;                 // 4 key words of 64 bits
constexpr int kStateEpoch = 0;                    // state words
constexpr int kStateTiles = 1;                    // [1, 1 + kMaxPasses)
constexpr int kStateCounts = 64;                  // [64, 64 + 256 * kMaxPasses)
constexpr unsigned kLookbackSpins = 1u << 20;
constexpr int kLookWidth = 16;                     // predecessor tiles read per look-back round

template <int PW> struct SortTile {
    static constexpr int kItems = PW <= 2 ? 8 : 4;                // keys per thread
    static constexpr int kTile = kSortThreads * kItems;          // keys per tile
    static constexpr int kWaveRun = kTile / kSortWaves;          // consecutive keys per wave
};

__device__ __forceinline__ unsigned status_tag(uint64_t v) { return (unsigned)(v >> 32); }
__device__ __forceinline__ bool status_incl(uint64_t v) { return (v >> 31) & 1u; }
__device__ __forceinline__ unsigned status_count(uint64_t v) { return (unsigned)v & 0x7fffffffu; }

__global__ __launch_bounds__(kSortThreads) void k_sort_init(uint32_t *__restrict__ state, int npass) {
    if (threadIdx.x == 0) state[kStateEpoch] = state[kStateEpoch] + 1;
    if (threadIdx.x < kMaxPasses) state[kStateTiles + threadIdx.x] = 0;
    for (int i = threadIdx.x; i < 256 * npass; i += kSortThreads) state[kStateCounts + i] = 0;
}

// Digit counts of all passes.  keys: word-major, PW words, stride m_max.
template <int PW>
__global__ __launch_bounds__(kSortThreads) void k_sort_counts(const uint64_t *__restrict__ keys, int stride,
                                                             const Counters *ctr, int npass,
                                                             uint32_t *__restrict__ state) {
    __shared__ uint32_t h[kMaxPasses * 256];
    const unsigned M = ctr->M;
    for (int i = threadIdx.x; i < 256 * npass; i += kSortThreads) h[i] = 0;
    __syncthreads();
    for (unsigned i = blockIdx.x * kSortThreads + threadIdx.x; i < M; i += gridDim.x * kSortThreads) {
#pragma unroll
        for (int w = 0; w < PW; w++) {
            const uint64_t k = keys[(size_t)w * stride + i];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int p = w * 8 + b;
                if (p < npass) atomicAdd(&h[p * 256 + ((k >> (8 * b)) & 255)], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 256 * npass; i += kSortThreads)
        if (h[i]) atomicAdd(&state[kStateCounts + i], h[i]);
}

// Exclusive scan of one value per thread over the workgroup's 256 threads.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *tmp /* kSortWaves */) {
    const int lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) tmp[wave] = x;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) before += (w < wave) ? tmp[w] : 0u;
    __syncthreads();
    return before + x - v;
}

// Reduce-then-scan form of a pass (many tiles): per-tile digit counts
// (tile-major hist[tile * 256 + digit]), then a column scan of that matrix in
// three launches over groups of `gt` tiles: group sums, a scan of the group
// sums (one workgroup per digit), and the running totals written back in
// place, so hist[tile][digit] becomes the output index of the tile's first
// key with that digit.  Group sums follow the tile counts in `hist`.
template <int PW>
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t *__restrict__ kin, int stride,
                                                           const Counters *ctr, int pass,
                                                           uint32_t *__restrict__ hist) {
    using T = SortTile<PW>;
    __shared__ uint32_t h[256];
    const unsigned M = ctr->M;
    const int tile = (int)blockIdx.x;
    if ((unsigned)tile * T::kTile >= M) return;
    h[threadIdx.x] = 0;
    __syncthreads();
    const int word = pass / 8, shift = (pass % 8) * 8;
#pragma unroll
    for (int r = 0; r < T::kItems; r++) {
        const unsigned i = (unsigned)tile * T::kTile + r * kSortThreads + threadIdx.x;
        if (i < M) atomicAdd(&h[(kin[(size_t)word * stride + i] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[(size_t)tile * 256 + threadIdx.x] = h[threadIdx.x];
}

template <int PW>
__global__ __launch_bounds__(kSortThreads) void k_sort_group_sums(const Counters *ctr, uint32_t *__restrict__ hist,
                                                                 int max_tiles, int gt) {
    using T = SortTile<PW>;
    const int ntiles = (int)((ctr->M + T::kTile - 1) / T::kTile);
    const int g = (int)blockIdx.x, t0 = g * gt;
    if (t0 >= ntiles) return;
    const int t1 = min(ntiles, t0 + gt);
    uint32_t s = 0;
    for (int t = t0; t < t1; t += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = t + q < t1 ? hist[(size_t)(t + q) * 256 + threadIdx.x] : 0u;
#pragma unroll
        for (int q = 0; q < 8; q++) s += v[q];
    }
    hist[(size_t)max_tiles * 256 + (size_t)g * 256 + threadIdx.x] = s;
}

// One workgroup per digit: exclusive scan of that digit's group sums (at most
// kSortThreads * 4 groups), offset by the digit's global start.
template <int PW>
__global__ __launch_bounds__(kSortThreads) void k_sort_group_scan(const Counters *ctr, int pass,
                                                                 const uint32_t *__restrict__ state,
                                                                 uint32_t *__restrict__ hist, int max_tiles,
                                                                 int gt) {
    using T = SortTile<PW>;
    __shared__ uint32_t tmp[kSortWaves];
    __shared__ uint32_t s_goff;
    const int d = (int)blockIdx.x;
    const int ntiles = (int)((ctr->M + T::kTile - 1) / T::kTile);
    const int ng = (ntiles + gt - 1) / gt;
    // the digit's global start: digits below it over all keys
    uint32_t below = 0;
    for (int e = threadIdx.x; e < d; e += kSortThreads) below += state[kStateCounts + pass * 256 + e];
    below = block_excl_scan(below, tmp) + below;   // inclusive: the last thread holds the total
    if (threadIdx.x == kSortThreads - 1) s_goff = below;
    __syncthreads();
    uint32_t *gs = hist + (size_t)max_tiles * 256 + d;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int g = threadIdx.x * 4 + q;
        v[q] = g < ng ? gs[(size_t)g * 256] : 0u;
        s += v[q];
    }
    uint32_t run = s_goff + block_excl_scan(s, tmp);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int g = threadIdx.x * 4 + q;
        if (g < ng) gs[(size_t)g * 256] = run;
        run += v[q];
    }
}

template <int PW>
__global__ __launch_bounds__(kSortThreads) void k_sort_tile_offsets(const Counters *ctr, uint32_t *__restrict__ hist,
                                                                   int max_tiles, int gt) {
    using T = SortTile<PW>;
    const int ntiles = (int)((ctr->M + T::kTile - 1) / T::kTile);
    const int g = (int)blockIdx.x, t0 = g * gt;
    if (t0 >= ntiles) return;
    const int t1 = min(ntiles, t0 + gt);
    uint32_t run = hist[(size_t)max_tiles * 256 + (size_t)g * 256 + threadIdx.x];
    for (int t = t0; t < t1; t += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = t + q < t1 ? hist[(size_t)(t + q) * 256 + threadIdx.x] : 0u;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (t + q < t1) hist[(size_t)(t + q) * 256 + threadIdx.x] = run;
            run += v[q];
        }
    }
}

// LOOKBACK: decoupled look-back over the tiles' status words (few tiles);
// otherwise the chunk totals and tile counts of the reduce-then-scan form.
template <int PW, bool LOOKBACK>
__global__ __launch_bounds__(kSortThreads) void k_sort_pass(const uint64_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ pin,
                                                           uint64_t *__restrict__ kout, uint32_t *__restrict__ pout,
                                                           int stride, Counters *ctr, int pass,
                                                           uint32_t *__restrict__ state,
                                                           uint64_t *__restrict__ status,
                                                           const uint32_t *__restrict__ hist, int max_tiles) {
    using T = SortTile<PW>;
    constexpr int ITEMS = T::kItems, TILE = T::kTile, RUN = T::kWaveRun;
    __shared__ uint64_t skey[TILE * PW];
    __shared__ uint32_t spay[TILE];
    __shared__ uint32_t whist[kSortWaves][256];
    __shared__ uint32_t gbase[256];      // output index of the tile's first key of each digit
    __shared__ uint32_t dstart[256];     // first LDS slot of each digit in the ordered tile
    __shared__ uint32_t tmp[kSortWaves];

    const unsigned M = ctr->M;
    const int ntiles = (int)((M + TILE - 1) / TILE);
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid / kWave;
    const int tile = (int)blockIdx.x;
    if (tile >= ntiles) return;
    for (int w = 0; w < kSortWaves; w++) whist[w][tid] = 0;
    __syncthreads();
    const int word = pass / 8, shift = (pass % 8) * 8;
    const unsigned tag = state[kStateEpoch] * (unsigned)kMaxPasses + (unsigned)pass;

    // Rank: wave `wave` owns keys [tile*TILE + wave*RUN, +RUN), 64 per round.
    uint64_t key[ITEMS][PW];
    uint32_t pay[ITEMS];
    int dg[ITEMS], rk[ITEMS];
    const unsigned base = (unsigned)tile * TILE + wave * RUN;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const unsigned i = base + r * kWave + lane;
        const bool valid = i < M;
#pragma unroll
        for (int w = 0; w < PW; w++) key[r][w] = valid ? kin[(size_t)w * stride + i] : 0ull;
        pay[r] = valid ? pin[i] : 0u;
        const uint32_t d = valid ? (uint32_t)((key[r][word] >> shift) & 255) : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = valid ? whist[wave][d] : 0u;
        const int below = __popcll(peers & lanemask_lt());
        if (valid && below == 0) whist[wave][d] = before + (uint32_t)__popcll(peers);
        dg[r] = valid ? (int)d : -1;
        rk[r] = (int)before + below;
    }
    __syncthreads();

    // Per digit (thread tid = digit): wave offsets, tile count, look-back.
    uint32_t cnt = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
        const uint32_t c = whist[w][tid];
        whist[w][tid] = cnt;   // now the wave's offset inside the digit
        cnt += c;
    }
    uint64_t *my = status + (size_t)tile * 256 + tid;
    const uint64_t hi = (uint64_t)tag << 32;
    uint32_t excl = 0;
    if constexpr (!LOOKBACK) {
        excl = hist[(size_t)tile * 256 + tid];   // output index of the digit's first key (global start included)
    } else if (tile == 0) {
        __hip_atomic_store(my, hi | (1ull << 31) | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(my, hi | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // Look back kLookWidth tiles per round (independent loads in flight),
        // taking counts in order until a tile with its running total; stop at
        // the first tile not yet published and poll again from there.
        bool timed_out = false;
        int j = tile - 1;
        unsigned spins = 0;
        while (j >= 0) {
            uint64_t v[kLookWidth];
#pragma unroll
            for (int q = 0; q < kLookWidth; q++)
                v[q] = j - q >= 0 ? __hip_atomic_load(status + (size_t)(j - q) * 256 + tid, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0ull;
            bool done = false, stalled = false;
#pragma unroll
            for (int q = 0; q < kLookWidth; q++) {
                if (done || stalled || j - q < 0) continue;
                if (status_tag(v[q]) != tag) {
                    stalled = true;
                    j -= q;   // resume at the unpublished tile
                    continue;
                }
                excl += status_count(v[q]);
                if (status_incl(v[q])) done = true;
            }
            if (done) break;
            if (!stalled) {
                j -= kLookWidth;
                continue;
            }
            if (++spins >= kLookbackSpins) {
                timed_out = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            // after one look-back has given up, the others stop waiting too
            if ((spins & 1023u) == 0 &&
                (__hip_atomic_load(&ctr->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kFlagSort)) {
                timed_out = true;
                break;
            }
        }
        if (timed_out) atomicOr(&ctr->flags, kFlagSort);
        __hip_atomic_store(my, hi | (1ull << 31) | (excl + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // Global offset of the digit: digits before it over all keys, plus this
    // digit's keys in earlier tiles (the reduce-then-scan totals include the
    // former already).
    const uint32_t goff = LOOKBACK ? block_excl_scan(state[kStateCounts + pass * 256 + tid], tmp) : 0u;
    const uint32_t ds = block_excl_scan(cnt, tmp);
    gbase[tid] = goff + excl;
    dstart[tid] = ds;
    __syncthreads();

    // Order the tile by digit in LDS, then write digit runs.
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        if (dg[r] < 0) continue;
        const uint32_t l = dstart[dg[r]] + whist[wave][dg[r]] + (uint32_t)rk[r];
#pragma unroll
        for (int w = 0; w < PW; w++) skey[w * TILE + l] = key[r][w];
        spay[l] = pay[r];
    }
    __syncthreads();
    const int len = (int)min((unsigned)TILE, M - (unsigned)tile * TILE);
    for (int l = tid; l < len; l += kSortThreads) {
        const uint64_t kw = skey[word * TILE + l];
        const int d = (int)((kw >> shift) & 255);
        const uint32_t o = gbase[d] + (uint32_t)l - dstart[d];
#pragma unroll
        for (int w = 0; w < PW; w++) kout[(size_t)w * stride + o] = skey[w * TILE + l];
        pout[o] = spay[l];
    }
}

size_t sort_state_words() { return kStateCounts + 256 * kMaxPasses; }

size_t sort_status_entries(int m_max, int pw) {
    const int tile = pw <= 2 ? SortTile<1>::kTile : SortTile<3>::kTile;
    const size_t tiles = ((size_t)(m_max > 0 ? m_max : 1) + tile - 1) / tile;
    return tiles * 256;
}

size_t sort_hist_words(int m_max, int pw) {
    const size_t tiles = sort_status_entries(m_max, pw) / 256;
    return (tiles + 4 * kSortThreads) * 256;   // tile counts, then up to 1024 group sums
}

template <int PW>
static void sort_launch(SortBufs &b, Counters *ctr, int m_max, int np, hipStream_t st) {
    using T = SortTile<PW>;
    const int tiles = (m_max + T::kTile - 1) / T::kTile;
    // groups of gt tiles (a multiple of 8), at most 4 * 256 groups
    const int gt = std::max(8, ((tiles + 4 * kSortThreads - 1) / (4 * kSortThreads) + 7) / 8 * 8);
    const int groups = (tiles + gt - 1) / gt;
    const int cblk = std::max(1, std::min(256, (m_max + kSortThreads * 16 - 1) / (kSortThreads * 16)));
    const bool lookback = b.lookback_tiles < 0 ? false : tiles <= b.lookback_tiles;
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(kSortThreads), 0, st, b.state, np);
    hipLaunchKernelGGL((k_sort_counts<PW>), dim3(cblk), dim3(kSortThreads), 0, st, b.pk_a, m_max, ctr, np,
                       b.state);
    uint64_t *kin = b.pk_a, *kout = b.pk_b;
    uint32_t *pin = b.pay_a, *pout = b.pay_b;
    for (int p = 0; p < np; p++) {
        if (lookback) {
            hipLaunchKernelGGL((k_sort_pass<PW, true>), dim3(tiles), dim3(kSortThreads), 0, st, kin, pin, kout, pout,
                               m_max, ctr, p, b.state, b.status, b.hist, tiles);
        } else {
            hipLaunchKernelGGL((k_sort_hist<PW>), dim3(tiles), dim3(kSortThreads), 0, st, kin, m_max, ctr, p, b.hist);
            hipLaunchKernelGGL((k_sort_group_sums<PW>), dim3(groups), dim3(kSortThreads), 0, st, ctr, b.hist, tiles,
                               gt);
            hipLaunchKernelGGL((k_sort_group_scan<PW>), dim3(256), dim3(kSortThreads), 0, st, ctr, p, b.state, b.hist,
                               tiles, gt);
            hipLaunchKernelGGL((k_sort_tile_offsets<PW>), dim3(groups), dim3(kSortThreads), 0, st, ctr, b.hist, tiles,
                               gt);
            hipLaunchKernelGGL((k_sort_pass<PW, false>), dim3(tiles), dim3(kSortThreads), 0, st, kin, pin, kout,
                               pout, m_max, ctr, p, b.state, b.status, b.hist, tiles);
        }
        std::swap(kin, kout);
        std::swap(pin, pout);
    }
}

hipError_t radix_sort(SortBufs &b, Counters *ctr, int m_max, int pw, int bits, bool *result_in_b, int *passes,
                      hipStream_t st) {
    *result_in_b = false;
    *passes = 0;
    if (m_max <= 1 || bits <= 0) return hipSuccess;
    const int np = (bits + 7) / 8;
    if (pw < 1 || pw > 4 || np > kMaxPasses || np > 8 * pw) return hipErrorInvalidValue;
    switch (pw) {
    case 1: sort_launch<1>(b, ctr, m_max, np, st); break;
    case 2: sort_launch<2>(b, ctr, m_max, np, st); break;
    case 3: sort_launch<3>(b, ctr, m_max, np, st); break;
    default: sort_launch<4>(b, ctr, m_max, np, st); break;
    }
    *result_in_b = (np % 2) == 1;
    *passes = np;
    return hipGetLastError();
}

} // namespace gt
