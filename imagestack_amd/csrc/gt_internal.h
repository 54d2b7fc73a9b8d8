// Host-side declarations shared by the kernel translation units and the
// C-ABI implementation.  Not part of the public interface (include/gtlattice.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gt {

constexpr int kMaxDH = 16;

// Device counters written by the splat/compaction kernels and read back by
// the host once per transform (one small D2H copy).
struct Counters {
    unsigned int created;      // vertices created in the table
    unsigned int flags;        // kFlag* bits
    unsigned int listed;       // slots appended by compaction (== M)
    unsigned int done_pixels;  // pixels fully splatted (used to extrapolate M on regrowth)
    int kmin[kMaxDH];          // per-coordinate min / max of the keys (biased uint16)
    int kmax[kMaxDH];
    unsigned long long probes;
    unsigned long long inserts;
    unsigned long long atomics;
};
constexpr unsigned kFlagFull = 1u;   // table reached its fill limit: regrow and redo
constexpr unsigned kFlagBad = 2u;    // NaN/Inf position or key outside int16

// Where pixel positions and values come from.
struct PixSrc {
    // GaussTransform form: d position planes, scaled by inv_sigma.
    const float *pos;
    int64_t pos_cs;
    float inv[kMaxDH];
    // JointBilateral form: Cr guide planes * ic, then x*ix, y*iy, t*it for the
    // filtered axes, values [im, 1].
    const float *ref;
    int64_t ref_cs;
    int cr;
    int W, H;
    int fx, fy, ft;
    float ic, ix, iy, it;
    // values: nval planes, plus a constant-1 channel when homog
    const float *val;
    int64_t val_cs;
    int nval;
    int homog;
};

struct LatConst {
    float sf[kMaxDH];
};

// Output of the slice.
struct OutDst {
    float *out;
    int64_t out_cs;
    int normalize;  // 1: write vd-1 channels divided by the last one
};

struct TableArgs {
    uint64_t *tkeys;   // capacity * KW words
    float *tvals;      // capacity * vd FP32 accumulators
    uint32_t mask;     // capacity - 1
    uint32_t max_fill; // regrow threshold on created vertices
    int *slot_id;      // capacity: slot -> sorted vertex id
};

// Launchers (gt_kernels.hip).  All are asynchronous on `st`.
hipError_t launch_clear(uint64_t *tkeys, size_t nwords, float *tvals, size_t nvals,
                        Counters *ctr, hipStream_t st);
hipError_t launch_splat(int d, bool bilat, bool dedup, const PixSrc &src, int n, int row_w,
                        int tile_w, int block, int vd, const LatConst &lc, const TableArgs &ta,
                        Counters *ctr, bool count, hipStream_t st);
hipError_t launch_compact(int d, const TableArgs &ta, uint32_t capacity, uint32_t *slot_list,
                          Counters *ctr, hipStream_t st);
hipError_t launch_pack(int d, const TableArgs &ta, const uint32_t *slot_list, int M,
                       const int *umin, const int *bits, int pw, uint64_t *pk, hipStream_t st);
hipError_t launch_finalize(int d, int storage, const TableArgs &ta, const uint32_t *sorted_slots,
                           int M, int vd, int vs, uint64_t *vkeys, void *vals, float *splat_copy,
                           hipStream_t st);
hipError_t launch_neighbors(int d, const TableArgs &ta, const uint64_t *vkeys, int M,
                            int2 *nbr, Counters *ctr, bool count, hipStream_t st);
hipError_t launch_blur(int storage, int vs, const int2 *nbr, int M, const void *in, void *out,
                       hipStream_t st);
hipError_t launch_slice(int d, bool bilat, int storage, const PixSrc &src, int n, int vd, int vs,
                        const LatConst &lc, const TableArgs &ta, const void *vals,
                        const OutDst &dst, Counters *ctr, bool count, hipStream_t st);
hipError_t launch_unpack_keys(int d, const uint64_t *vkeys, int M, int16_t *keys, hipStream_t st);
hipError_t launch_store_to_f32(int storage, const void *vals, int M, int vd, int vs, float *out,
                               hipStream_t st);

// Radix sort (gt_sort.hip): sorts M items with pw 64-bit key words stored
// word-major (pk[w*M + i], word 0 least significant) and a uint32 payload, by
// the low `bits` bits.  Stable LSD passes of 8 bits.  Buffers a/b are
// ping-pong; returns in *result_in_b whether the sorted data ended in b.
struct SortBufs {
    uint64_t *pk_a, *pk_b;
    uint32_t *pay_a, *pay_b;
    uint32_t *hist;        // 256 * nblocks
    size_t hist_cap;       // entries available in hist
};
size_t radix_hist_entries(int M);
hipError_t radix_sort(SortBufs &b, int M, int pw, int bits, bool *result_in_b, int *passes,
                      hipStream_t st);

} // namespace gt
