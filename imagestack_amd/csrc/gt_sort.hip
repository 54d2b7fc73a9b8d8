// LSD radix sort of (multi-word key, uint32 payload) pairs for the lattice
// vertex keys, hand-written for gfx950 (wave64).  Each 8-bit pass is
// histogram -> exclusive scan -> stable scatter.  Stability inside a
// workgroup comes from processing items in index order, 256 at a time, and
// ranking equal digits inside each wave with eight ballots (no LDS sort).
// The item count lives on the device (Counters::M): grids are sized for the
// host's upper bound and idle workgroups exit at once, so no host round trip
// sits between the compaction and the sort.
#include "gt_device.h"
#include "gt_internal.h"

#include <utility>

namespace gt {

constexpr int kSortThreads = 256;
constexpr int kSortTiles = 16;                        // 256-item tiles per workgroup
constexpr int kSortItems = kSortThreads * kSortTiles; // items per workgroup

__device__ __forceinline__ int sort_blocks(unsigned M) { return (int)((M + kSortItems - 1) / kSortItems); }

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t *__restrict__ kw, const Counters *ctr,
                                                            int shift, uint32_t *__restrict__ hist) {
    const unsigned M = ctr->M;
    const int nblk = sort_blocks(M);
    if ((int)blockIdx.x >= nblk) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const unsigned base = blockIdx.x * kSortItems;
#pragma unroll 4
    for (int t = 0; t < kSortTiles; t++) {
        const unsigned i = base + t * kSortThreads + threadIdx.x;
        if (i < M) atomicAdd(&h[(kw[i] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit-major histogram (256 * nblk entries, a multiple
// of 4) in one workgroup of 1024 threads: each thread owns a contiguous run of
// uint4s, loaded four at a time so the loads overlap.
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t *__restrict__ hist, const Counters *ctr) {
    __shared__ uint32_t part[1024];
    const int E4 = 64 * sort_blocks(ctr->M);
    const int chunk = (E4 + 1023) / 1024;
    const int lo = threadIdx.x * chunk;
    const int hi = min(E4, lo + chunk);
    uint4 *h4 = reinterpret_cast<uint4 *>(hist);
    uint32_t s = 0;
#pragma unroll 4
    for (int i = lo; i < hi; i++) {
        const uint4 v = h4[i];
        s += v.x + v.y + v.z + v.w;
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
#pragma unroll 4
    for (int i = lo; i < hi; i++) {
        const uint4 v = h4[i];
        uint4 o;
        o.x = run; run += v.x;
        o.y = run; run += v.y;
        o.z = run; run += v.z;
        o.w = run; run += v.w;
        h4[i] = o;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t *__restrict__ pk_in,
                                                               const uint32_t *__restrict__ pay_in,
                                                               uint64_t *__restrict__ pk_out,
                                                               uint32_t *__restrict__ pay_out, const Counters *ctr,
                                                               int stride, int pw, int word, int shift,
                                                               const uint32_t *__restrict__ hist) {
    constexpr int NW = kSortThreads / kWave;
    const unsigned M = ctr->M;
    const int nblk = sort_blocks(M);
    if ((int)blockIdx.x >= nblk) return;
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[NW][256];
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    base[tid] = hist[(size_t)tid * nblk + blockIdx.x];
    const unsigned b0 = blockIdx.x * kSortItems;
    for (int t = 0; t < kSortTiles; t++) {
#pragma unroll
        for (int w = 0; w < NW; w++) wcnt[w][tid] = 0;
        __syncthreads();
        const unsigned i = b0 + t * kSortThreads + tid;
        const bool valid = i < M;
        const uint32_t dg = valid ? (uint32_t)((pk_in[(size_t)word * stride + i] >> shift) & 255) : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & lanemask_lt());
        if (valid && rank == 0) wcnt[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[dg] + (uint32_t)rank;
#pragma unroll
            for (int w = 0; w < NW; w++) off += (w < wave) ? wcnt[w][dg] : 0u;
            for (int w = 0; w < pw; w++) pk_out[(size_t)w * stride + off] = pk_in[(size_t)w * stride + i];
            pay_out[off] = pay_in[i];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) add += wcnt[w][tid];
        base[tid] += add;
        __syncthreads();
    }
}

size_t radix_hist_entries(int m_max) {
    const size_t nblk = ((size_t)m_max + kSortItems - 1) / kSortItems;
    return 256 * (nblk ? nblk : 1);
}

hipError_t radix_sort(SortBufs &b, const Counters *ctr, int m_max, int pw, int bits, bool *result_in_b, int *passes,
                      hipStream_t st) {
    *result_in_b = false;
    *passes = 0;
    if (m_max <= 1 || bits <= 0) return hipSuccess;
    const int nblk = (m_max + kSortItems - 1) / kSortItems;
    if ((size_t)256 * nblk > b.hist_cap) return hipErrorInvalidValue;
    const int np = (bits + 7) / 8;
    uint64_t *pin = b.pk_a, *pout = b.pk_b;
    uint32_t *yin = b.pay_a, *yout = b.pay_b;
    for (int p = 0; p < np; p++) {
        const int word = p / 8, shift = (p % 8) * 8;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kSortThreads), 0, st, pin + (size_t)word * m_max, ctr, shift,
                           b.hist);
        hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, st, b.hist, ctr);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kSortThreads), 0, st, pin, yin, pout, yout, ctr, m_max,
                           pw, word, shift, b.hist);
        std::swap(pin, pout);
        std::swap(yin, yout);
    }
    *result_in_b = (np % 2) == 1;
    *passes = np;
    return hipGetLastError();
}

} // namespace gt
