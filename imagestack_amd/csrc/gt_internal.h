// Host/device declarations shared by the kernel translation units and the
// C-ABI implementation.  Not part of the public interface (include/gtlattice.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gt_keyplan.h"

struct gt_ctx;             // the C ABI's opaque context (gt_api.cpp)
struct gt_density_table;   // vertex densities shared by contexts (gt_api.cpp)

namespace gt {

// Device counters, read back by the host once per transform.
struct Counters {
    unsigned int M;            // lattice vertices (written by the compaction scan)
    unsigned int flags;        // kFlag* bits
    unsigned int filled;       // occupied table slots at the compaction (also after a full table)
    unsigned int pad;
    unsigned long long probes;   // profiling level 2 only
    unsigned long long inserts;
    unsigned long long atomics;
};
constexpr int kCompactSlots = 2048;   // table slots per compaction workgroup (chunk)
constexpr unsigned kFlagFull = 1u;    // an insert exceeded the probe limit: regrow and redo
constexpr unsigned kFlagBad = 2u;     // NaN/Inf position or key outside int16
constexpr unsigned kFlagRange = 4u;   // a key outside the packing plan: replan and redo
constexpr unsigned kFlagSort = 8u;    // a radix-sort look-back gave up waiting (a bug, never expected)
constexpr unsigned kFlagNonFinite = 16u;   // the values hold NaN/Inf (marks in TableArgs::tnf, not an error)
constexpr unsigned kFlagWait = 32u;   // a table key word awaited from its writer never came (a bug, never expected)

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
    int W, H, F;     // image size (F frames: rows t * H + y)
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

// A batch of independent joint-bilateral images filtered through one lattice
// (the image index is the key's tag field).  Frames are 1; all images share
// the channel counts and storage.
struct BatchImage {
    const float *ref;   // Cr guide planes of W*H pixels
    const float *im;    // C image planes
    float *out;         // C output planes
    int W, H;
    float ic, ix, iy;   // 1/sigma_r, 1/sigma_s, 1/sigma_s
    int tile0;          // first tile of this image in the batch's tile order
    // Table parts of a batch: this image's keys take the slice [start24,
    // start24 + its vertex share) of [0, 2^24) -- a row band of it by their
    // fields' sum u in [0, part_range]: start24 + (u * part_mulw >> 24) --
    // then scaled to the call's P parts.
    uint32_t part_start24;
    int part_range;
    uint64_t part_mulw;
};

struct BatchArgs {
    const BatchImage *img;  // device array of `count`
    int count;
};

// Output of the slice.
struct OutDst {
    float *out;
    int64_t out_cs;
    int normalize;  // 1: write vd-1 channels divided by the last one
};

struct TableArgs {
    uint64_t *tkeys;               // capacity * W words
    unsigned long long *tacc;      // capacity * vd fixed-point accumulators (two's complement)
    uint32_t *tnf;                 // capacity: NaN/Inf marks of each slot's sums, 2 bits per channel
                                   // (bit 2c: a +Inf contribution, bit 2c+1: a -Inf one; both: NaN);
                                   // zero between calls (the finalize clears what it reads)
    uint32_t mask;                 // capacity - 1
    int *slot_id;                  // capacity: slot -> vertex id (idshift == 0)
    int idshift;                   // > 0 (one-word keys with room to spare): after the finalize a
                                   // table word holds (id << idshift) | key, so a lookup reads the
                                   // vertex id with the key; slot_id is not used
    uint64_t keymask;              // (1 << idshift) - 1
    // part_mul != 0 (a one-frame bilateral form over image rows): the
    // table is P = 2^part_lg parts of capacity / P slots; a key's home slot
    // lies in part ((c_0 + ... + c_{D-1} + part_base) clamped to [0,
    // part_range]) * part_mul >> 24, its hash picking the slot inside the
    // part.  The sum is -c_D, which follows the last position channel (the
    // row) alone, so each horizontal band of the image keeps its vertices in
    // one part of the table.
    uint32_t part_mul;
    int part_base, part_range;
    const BatchImage *part_img;    // a batch: per-image part slices, by the key's image tag
    int part_lg;                   // log2 of the part count P (>= 8)
    int part_shift;                // log2(capacity / P)
    // capacity bytes or null: after the finalize (ids in compaction order),
    // slot s's tag -- 0 empty, else 0x80 | 7 bits of its key's hash
    // (key_tag8).  A lookup probes the tags and reads key words only where the
    // tag matches: a lattice of wide keys is gigabytes of key words, and most
    // neighbour lookups of a sparse one miss, so the probes stay in the tags of
    // the XCD's own table parts, in its L2.
    uint8_t *tags;
};

// Replay record of the LDS-tiled splat, so that a slice at the splat's own
// positions needs neither the vertex keys nor the tile table's inserts again
// (the reference's replay slice, src/Permutohedral.h:405-427, keeps an
// (offset, weight) pair per pixel and vertex the same way; here the slice
// recomputes the weights, which costs less than moving them through HBM).
// Entry q = tile * 256 + thread of the splat's workgroup for that tile; the
// slice runs the same tile decomposition.  Entry q is the replay_dwords(D)
// words rec[q * that ...]: for vertex j in the splat's per-lane rotated order
// (vertex (j + thread % (D+1)) % (D+1)), the 10-bit field j % 3 of word j / 3
// holds its slot in the tile's LDS table, or kReplayGlobal when the tile
// table had no room for it and the splat went to the global table directly,
// at slot gs[j * stride + q].  Per tile t: tcnt[t] distinct keys, entry k of
// tent[t * kTileKeys + k] = {LDS slot, global table slot}.  rec == nullptr:
// no record.
struct Replay {
    uint32_t *rec;
    uint32_t *gs;
    uint2 *tent;
    uint32_t *tcnt;
    int64_t stride;   // entries: tiles * 256
};
constexpr unsigned kReplayGlobal = 1023u;
// Words of one replay entry: ceil((D+1)/3) slot words as one 8- or 16-byte
// vector (2 words up to d = 5).
constexpr int replay_dwords(int D) { return (D + 3) / 3 <= 2 ? 2 : 4; }
constexpr int kTileKeys = 512;   // LDS tile table slots of the tiled splat and slice
constexpr int kTilePixels = 256;   // pixels of a tile: one workgroup of the tiled kernels

// Tiles of an n-pixel image with rows of row_w pixels cut into tile_w x
// (256 / tile_w) tiles (tile_w a divisor of 256); when that does not apply,
// tile_w becomes 0 and the tiles are runs of 256 pixels.
inline int64_t pixel_tiles(int n, int row_w, int &tile_w) {
    if (tile_w > 0 && row_w > 0 && (kTilePixels % tile_w) == 0 && tile_w <= kTilePixels) {
        const int tiles_x = (row_w + tile_w - 1) / tile_w;
        const int th = kTilePixels / tile_w;
        const int rows = n / row_w;
        return (int64_t)tiles_x * ((rows + th - 1) / th);
    }
    tile_w = 0;
    return ((int64_t)n + kTilePixels - 1) / kTilePixels;
}

// Vertex ids grouped by eighths of a partitioned table (ids in compaction
// order): eighth x holds ids [off[x * cpp], off[(x + 1) * cpp]) (M for
// x = 8), off being the compaction's per-chunk offsets.  cpp = 0: none.
struct PartMap {
    const uint32_t *off;
    int cpp;
};

// Fixed-point splat accumulation: channel c of a weighted value is added as
// round(x * s[c]) to 64-bit integers, s[c] a power of two chosen so that no
// vertex sum can exceed 2^62 in magnitude (|value| <= the channel's largest
// finite magnitude in this call, weights <= 1, at most n contributions).
// Integer sums are exact and independent of the order of the atomics;
// finalize converts them back with inv_s[c].  The lattice pipeline computes
// it on the device at the start of every call (k_clear_scale), so the scales
// depend on the call's own values only.  nonfinite: some value is NaN/Inf;
// those contributions are kept as marks (TableArgs::tnf), not sums.
//
// BF16/FP16 value storage keeps each channel scaled by a power of two: the
// values written by blur pass j are true * 2^e_j with e_j = 14 -
// ceil(log2(m_j)), m_j the largest finite magnitude of the pass's input
// (vmax_bits[j], as float bits: [0] the splat sums, [j] pass j-1's output).
// The blur never raises the largest magnitude (its weights sum to at most 1),
// so stored values stay below 2^14 and the smallest ones keep FP16's normal
// range 28 binades further down, where d + 1 unscaled passes over a sparse
// lattice would halve them into subnormals.  Powers of two scale exactly, so
// the only rounding is the storage conversion; the slice (and an export)
// undo e_D with the same exactness.
struct ValScale {
    float s[kMaxDH];
    float inv_s[kMaxDH];
    unsigned nonfinite;
    unsigned vmax_bits[kMaxDH + 2][kMaxDH];
};

// Launchers (gt_launch_*.hip, gt_sort.hip).  All are asynchronous on `st`.
// Kernels after the compaction read the vertex count M from ctr->M on the
// device; `m_max` only sizes their grids.
// Per-plane min, max and a non-finite indicator (1.0 if any NaN/Inf) of two
// plane groups (A then B, n pixels each) -> out[3*plane..].
hipError_t launch_minmax(const float *pa, int64_t csa, int na, const float *pb, int64_t csb, int nb, int n,
                         float *partial, int nblk, float *out, hipStream_t st);
hipError_t launch_bounds_batch(const BatchArgs &ba, int ntiles, int tile_w, int Cr, int C, float *partial, int nblk,
                               float *out, hipStream_t st);
// What a lattice call clears before its splat: nwords table words to
// kEmpty, nacc sums and nmarks non-finite marks to zero.
struct ClearArgs {
    uint64_t *tkeys;
    size_t nwords;
    unsigned long long *tacc;
    size_t nacc;
    uint32_t *tnf;
    size_t nmarks;
    uint8_t *tags;   // the per-slot tags (nullptr: none), ntags bytes, a multiple of 4
    size_t ntags;
};
// The start of a lattice call in two launches: the largest finite magnitude
// and a non-finite indicator of each of the nval value planes (n pixels
// each); then the table cleared (ClearArgs), the counters zeroed and this
// call's fixed-point scales for vd channels (a homogeneous channel of ones
// after the planes when homog) written to the device ValScale, kFlagNonFinite
// raised in ctr when a value is NaN or Inf.  partial holds nblk * nval * 3
// floats.
hipError_t launch_value_scale(const ClearArgs &cl, const float *val, int64_t cs, int nval, int homog, int n, int vd,
                              float *partial, int nblk, ValScale *out, Counters *ctr, hipStream_t st);
// The same over the image planes of a batch (n: the batch's pixel count).
hipError_t launch_value_scale_batch(const ClearArgs &cl, const BatchArgs &ba, int ntiles, int tile_w, int C, int n,
                                    int vd, float *partial, int nblk, ValScale *out, Counters *ctr, hipStream_t st);
// rp: the tiled splat's replay record (rp.rec == nullptr: none; the direct
// splat writes none).
hipError_t launch_splat(int d, bool packed, bool bilat, bool dedup, const PixSrc &src, int n, int row_w, int tile_w,
                        int vd, const LatConst &lc, const KeyPlan &kp, const ValScale *vsc, const TableArgs &ta,
                        const Replay &rp, Counters *ctr, bool count, int wpc, hipStream_t st);
// Batched forms (bilateral, packed keys, d <= 6, vd <= 4): one workgroup per
// tile, ntiles tiles over all images.  wpc: the waves-per-CU cap of the tiled
// and neighbour kernels (0: none).
hipError_t launch_splat_batch(int d, const PixSrc &tmpl, const BatchArgs &ba, int ntiles, int tile_w, int vd,
                              const LatConst &lc, const KeyPlan &kp, const ValScale *vsc, const TableArgs &ta,
                              const Replay &rp, Counters *ctr, bool count, int wpc, hipStream_t st);
hipError_t launch_slice_batch(int d, int storage, const PixSrc &tmpl, const BatchArgs &ba, int ntiles, int tile_w,
                              int vd, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta, const void *vals,
                              const unsigned *sbits, const OutDst &tmpl_dst, Counters *ctr, bool count, int wpc,
                              hipStream_t st);
// Neighbour entries set to {-1, -1} with the finalize (nbr == nullptr: none).
struct NbrInit {
    int2 *nbr;
    int stride;   // entries between axes
    int axes;
    int grid;     // > 0: by a fill kernel after the finalize, its grid sized for this many vertices
                  // (it loops to M); 0: by the finalize's workgroups, each its chunk's ids
};

// A finalize that scans the compaction's chunk counts itself instead of a
// separate one-workgroup scan (tables of at most kSelfScanChunks chunks):
// each workgroup sums the counts of the chunks before its own; offs receives
// those offsets (the PartMap then reads them), wctr the vertex count M (0
// after a full table or over `limit`) and the filled count.  offs == nullptr:
// the counts were scanned already.
struct SelfScan {
    uint32_t *offs;
    Counters *wctr;
    int nchunks;
    unsigned limit;
};
constexpr int kSelfScanChunks = 2048;
// limit: the per-vertex buffers' size; a table holding more vertices counts as
// full (kFlagFull, M = 0).
hipError_t launch_compact(int words, const TableArgs &ta, uint32_t capacity, uint32_t *counts, uint32_t *slot_list,
                          uint64_t *sort_keys, bool packed, Counters *ctr, int limit, hipStream_t st);
hipError_t launch_pack(int d, const TableArgs &ta, const uint32_t *slot_list, const Counters *ctr, int m_max,
                       const KeyPlan &kp, int pw, uint64_t *pk, hipStream_t st);
// sorted_keys: the sorted packed keys (one word each), or null to gather the
// table's key words by slot.
hipError_t launch_finalize(int words, int storage, const TableArgs &ta, const ValScale *vsc, const uint32_t *sorted,
                           const uint64_t *sorted_keys, const Counters *ctr, int m_max, int vd, int vs,
                           uint64_t *vkeys, void *vals, float *splat_copy, hipStream_t st);
hipError_t launch_neighbors(int d, bool packed, const KeyPlan &kp, const uint64_t *vkeys, const Counters *ctr,
                            int m_max, int2 *nbr, int wpc, hipStream_t st);
// Compaction with the finalize fused in, for ids in compaction order (no
// sort, no slot list): count, scan, then k_write_final.
// offs: where the chunk offsets go when the table has at most
// kSelfScanChunks chunks (the finalize scans the counts itself); else they
// replace the counts in place.
// ninit.nbr != nullptr: the M vertices' neighbour entries are also set to
// {-1, -1} on each of ninit.axes axes (stride ninit.stride), for the
// neighbour form that fills the second half of each entry by scatter.
hipError_t launch_compact_final(int words, int storage, const TableArgs &ta, uint32_t capacity, uint32_t *counts,
                                uint32_t *offs, const ValScale *vsc, int vd, int vs, uint64_t *vkeys, void *vals,
                                Counters *ctr, int limit, const NbrInit &ninit, hipStream_t st);
// The same table by hash lookups in the complete global table, for vertex ids
// in compaction order (no sort).
// With pm.cpp > 0, workgroups on XCD x take the vertices of table part x.
// per_vertex: one thread per vertex for all axes (k_nbr_hash_vertex), else
// one per (vertex, axis) (k_nbr_hash, more threads for small lattices).
// m_max: the neighbour table's stride; m_grid (<= m_max): the vertices the
// grid is sized for (the kernel loops past it, so any M <= m_max is covered).
// half (per_vertex only): each vertex looks up its sign-0 neighbour u on
// every axis, writes it and writes itself as u's sign-1 neighbour (the
// relation is symmetric); the entries must hold {-1, -1} beforehand
// (launch_compact_final's ninit).
hipError_t launch_neighbors_hash(int d, bool packed, const KeyPlan &kp, const uint64_t *vkeys, const Counters *ctr,
                                 int m_max, int m_grid, const TableArgs &ta, const PartMap &pm, bool per_vertex,
                                 bool half, int2 *nbr, hipStream_t st);
// Blur pass `pass` (0..d): in holds FP32 for pass 0 of a narrow storage (the
// finalize writes FP32 there) and the storage type otherwise; vsc supplies the
// narrow storages' scales.
// m_max: the vertices the grid is sized for (the kernel loops past it).
hipError_t launch_blur(int storage, int vs, int pass, const ValScale *vsc, const int2 *nbr, const Counters *ctr,
                       int m_max, const PartMap &pm, const void *in, void *out, hipStream_t st);
// Passes j and j+1 in one launch (FP32 storage, vs <= 4): nbr_j, nbr_k their
// neighbour rows; the same result as two launch_blur calls.
hipError_t launch_blur2(int vs, const int2 *nbr_j, const int2 *nbr_k, const Counters *ctr, int m_max,
                        const PartMap &pm, const float *in, float *out, hipStream_t st);
// sbits: the scale source of a narrow storage's last blur pass
// (ValScale::vmax_bits[d]); ignored for FP32.
// Slice from the tiled splat's replay record (the splat's own positions,
// d <= 10, vd <= 4; batch: ba.count > 0, over ntiles tiles); the same tile
// decomposition as that splat (n, row_w, tile_w or the batch's tiles).
// hctr: host memory (pinned, device-visible) that receives the counters, or
// null.
hipError_t launch_slice_replay(int d, bool bilat, int storage, const PixSrc &src, int n, int row_w, int tile_w,
                               const BatchArgs &ba, int ntiles, int vd, const LatConst &lc, const KeyPlan &kp,
                               const TableArgs &ta, const Replay &rp, const void *vals, const unsigned *sbits,
                               const OutDst &dst, Counters *ctr, Counters *hctr, int wpc, hipStream_t st);
hipError_t launch_slice(int d, bool packed, bool bilat, bool tiled, int storage, const PixSrc &src, int n, int row_w,
                        int tile_w, int vd, int vs, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta,
                        const void *vals, const unsigned *sbits, const OutDst &dst, Counters *ctr, bool count,
                        int wpc, hipStream_t st);
hipError_t launch_unpack_keys(int d, bool packed, const KeyPlan &kp, const uint64_t *vkeys, int M, int16_t *keys,
                              hipStream_t st);
// First vertex of each image in a sorted batched lattice (keys ordered by
// their image tag): starts[i] for i in [0, count], starts[count] = M.
hipError_t launch_tag_starts(const uint64_t *vkeys, int M, int coord_bits, int count, uint32_t *starts,
                             hipStream_t st);
hipError_t launch_store_to_f32(int storage, const void *vals, const unsigned *sbits, int M, int vd, int vs,
                               float *out, hipStream_t st);
hipError_t launch_view_copy(const float *src, int64_t sys, int64_t sts, int64_t scs, float *dst, int64_t dys,
                            int64_t dts, int64_t dcs, int W, int H, int F, int C, hipStream_t st);

// Radix sort (gt_sort.hip): stable LSD sort of up to m_max items (the live
// count is ctr->M on the device) with `pw` key words stored word-major at
// stride m_max (pk[w*m_max + i], word 0 least significant) and a uint32
// payload, on the low `bits` bits, 8 bits per pass.  a/b are ping-pong
// buffers; *result_in_b tells where the sorted data ended.
struct SortBufs {
    uint64_t *pk_a, *pk_b;      // keys, word-major (word w of key i at w * m_max + i)
    uint32_t *pay_a, *pay_b;    // payloads
    uint32_t *state;            // sort_state_words(): epoch, tile counters, digit counts (zeroed once)
    uint64_t *status;           // sort_status_entries(): look-back words (zeroed once)
    uint32_t *hist;             // sort_hist_words(): tile counts and chunk sums of the reduce-then-scan form
    int lookback_tiles;         // decoupled look-back up to this many tiles, reduce-then-scan above (-1: never)
};
size_t sort_state_words();
size_t sort_status_entries(int m_max, int pw);
size_t sort_hist_words(int m_max, int pw);
// Bilateral grid (gt_grid.hip): at most kMaxGridD position dimensions.
constexpr int kMaxGridD = 6;
struct GridArgs {
    int vd;
    float sf;                     // sqrtf(1/3 + (taps - 1)/4), taps = 5
    float mn[kMaxGridD];          // minimum of position * sf per dimension
    int size[kMaxGridD];          // ceil(max - min) + 1 cells
    int64_t stride[kMaxGridD];    // in elements: stride[0] = vd
    uint32_t *marks;              // per element: the class of the NaN/Inf contributions (bit 0 +Inf, bit 1
                                  // -Inf, both NaN), or null when every value is finite
};
// Zero, splat, finish, 2 blur rounds per dimension, slice.  ev (optional):
// 6 events recorded at the stage boundaries (splat, -, -, blur, slice).
hipError_t launch_grid(int d, bool bilat, const PixSrc &splat, int n, int row_w, const PixSrc &slice, int n_slice,
                       const OutDst &dst, const GridArgs &g, int64_t cells, const ValScale &vsc,
                       unsigned long long *acc, float *a, float *b, hipEvent_t *ev, hipStream_t st);

// Patch projection (Convolve zero/inner with a K*C filter bank, K <= 16):
// filters [k][c][tap] are reordered into wt (K*C*fw*fh*ft floats), then
// out (K planes of W*H*F) is written.
hipError_t launch_patch_project(const float *im, int W, int H, int F, int C, const float *filters, int fw, int fh,
                                int ft, int K, float *wt, float *out, hipStream_t s);
size_t patch_lds_bytes(int fw, int fh);

// The context's device and current stream (for code outside gt_api.cpp).
int ctx_device(gt_ctx *c);
hipStream_t ctx_stream(gt_ctx *c);

// Sets the thread's gt_last_error() message; returns code.
int set_error(int code, const char *msg);
// True when an allocation is to fail (gt_debug_fail_allocations); every
// device or pinned allocation of the library asks first.
bool alloc_should_fail();
// Forgets the context's cached key-packing plan, so the next call bounds its
// inputs afresh (a server's next request is an unrelated image).
void invalidate_plan(::gt_ctx *c);
// A vertex-density table several contexts share (the serving loop's
// workers); it must outlive them.
::gt_density_table *density_table_new();
void density_table_delete(::gt_density_table *t);
void share_density(::gt_ctx *c, ::gt_density_table *t);
// Sorts the first ctr->M (key, payload) pairs by key (bits significant bits,
// pw words); the result is in the b buffers when *result_in_b.
hipError_t radix_sort(SortBufs &b, Counters *ctr, int m_max, int pw, int bits, bool *result_in_b, int *passes,
                      hipStream_t st);

} // namespace gt
