// C-ABI implementation of libgtlattice.so: context, device buffers, the
// splat -> sort -> blur -> slice pipeline, and the JointBilateral adapter.
//
// One call runs the whole pipeline on the context's stream with a single
// host synchronisation at the end (the vertex count M stays on the device).
// The end-of-call flags can ask for a redo: a table that filled up is grown,
// a key outside the packing plan triggers a new plan from fresh position
// bounds.  Host arithmetic that feeds the kernels (scale factors,
// reciprocals) is compiled with -ffp-contract=off and follows the
// reference's expressions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gt_internal.h"
#include "gt_plan.h"
#include "../../include/gtlattice.h"

namespace {

// The thread's last error message (a fixed buffer: reporting an error
// allocates nothing).
thread_local char g_err[512];

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define GT_HIP(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(GT_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_));     \
    } while (0)

#define GT_TRY(x)                     \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != GT_OK) return rc_; \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

uint32_t next_pow2(uint64_t v) {
    uint64_t c = 1024;
    while (c < v) c <<= 1;
    return c > (1ull << 31) ? (1u << 31) : (uint32_t)c;
}

// Radix sort: decoupled look-back while the tiles fit the GPU at once, the
// reduce-then-scan form beyond (the look-back's running totals advance a few
// tiles per status round trip, which bounds it once tiles queue up).
constexpr int kLookbackTiles = 256;

int value_stride(int vd) { return vd <= 2 ? 2 : vd <= 4 ? 4 : vd <= 8 ? 8 : 16; }

// Compaction chunks of a table of cap slots.
unsigned chunks_of(uint32_t cap) { return (cap + gt::kCompactSlots - 1) / gt::kCompactSlots; }
size_t store_bytes(int storage) { return storage == GT_STORE_F32 ? 4 : 2; }

} // namespace

namespace {
// gt_debug_fail_allocations: allocations to let through, then to fail
std::mutex g_fail_mu;
int g_fail_after = 0, g_fail_count = 0;
} // namespace

namespace gt {
int set_error(int code, const char *msg) { return fail(code, "%s", msg); }
bool alloc_should_fail() {
    std::lock_guard<std::mutex> lk(g_fail_mu);
    if (g_fail_count <= 0) return false;
    if (g_fail_after > 0) {
        g_fail_after--;
        return false;
    }
    g_fail_count--;
    return true;
}
} // namespace gt

extern "C" int gt_debug_fail_allocations(int after, int count) {
    std::lock_guard<std::mutex> lk(g_fail_mu);
    g_fail_after = after > 0 ? after : 0;
    g_fail_count = count > 0 ? count : 0;
    return GT_OK;
}

// Vertices per pixel seen for a filter (form, d, vd and the position
// scales): the table size of its next call at any image size.  A context
// keeps its own entries, or uses a table shared with other contexts (the
// serving loop's workers), locked because those run concurrently.
struct gt_density {
    uint64_t key = 0;
    float per_pixel = 0.f;
};
struct gt_density_table {
    std::mutex mu;
    gt_density e[32];
    int next = 0;
};

struct gt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr;
    int profiling = 0;   // 0 off, 1 HIP events, 2 events + probe/atomic counters
    hipEvent_t ev[8] = {};
    DevBuf tkeys, tacc, tnf, vscale, vs_part, ctr, counts, slot_list, slot_id, pk_a, pk_b, pay_b, hist, sort_status, sort_hist, vkeys,
        vals_a, vals_b, nbr, splat_copy, mm_part, mm_out;
    DevBuf stage_in, stage_out, view_a, view_b, view_o;
    DevBuf grid_acc, grid_a, grid_b;   // bilateral grid: fixed-point sums, FP32 ping-pong
    DevBuf grid_marks;                 // bilateral grid: NaN/Inf marks (calls with such values only)
    DevBuf patch_w, patch_guide;       // non-local means: reordered filters, projected guide
    DevBuf rp_rec, rp_gs, rp_tent, rp_tcnt;   // the tiled splat's replay record
    DevBuf tags;                               // per-slot key tags (hash ids, wide keys)
    gt::Counters *h_ctr = nullptr;   // pinned
    float *h_mm = nullptr;           // pinned, 3 * 2 * kMaxDH
    // capacity memory: last lattice size seen for (n, d), and vertices per
    // pixel seen for a filter (form, d, vd and the scales), used when the
    // same filter comes with a new image size
    int last_n = -1, last_d = -1;
    unsigned last_M = 0;
    size_t acc_zero = 0;   // leading bytes of tacc known to hold zeros
    bool tnf_clean = false;   // the non-finite marks are all zero (a complete call leaves them so)
    gt_density density[16];
    int density_next = 0;
    gt_density_table *shared_density = nullptr;   // not owned
    // packing plan cache (position-source kind, d): reused until a key falls outside it
    bool plan_valid = false;
    int plan_d = 0, plan_kind = -1, plan_vd = 0;
    int plan_n = 0;                  // pixels of the call the plan was made for
    float last_lo = 0, last_hi = 0;  // its last position channel's scaled bounds

    gt::KeyPlan plan{};
    // cached pipeline graph
    hipGraphExec_t gexec = nullptr;
    std::string gsig;
    int gpasses = 0;
    void *gvals = nullptr;
    // batch descriptors
    DevBuf batch_desc, tag_starts;
    // vertices of each image of the last batch whose lattice was sorted,
    // for the batch with these sizes and filters (the part slices of the next)
    uint64_t batch_sig = 0;
    std::vector<uint32_t> batch_verts;
    gt::BatchImage *h_batch = nullptr;   // pinned
    int h_batch_cap = 0;
    // lattice kept for export
    bool have = false;
    int M = 0, d = 0, vd = 0, vs = 0, storage = 0;
    bool packed = false;
    uint32_t cap = 0;
    void *vals_cur = nullptr;
};

namespace {

// Set while the no-allocation entry point runs on this thread: its buffers
// are fixed shares of the scratch array, and a buffer too small is an error.
thread_local bool g_no_alloc = false;

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return GT_OK;
    if (g_no_alloc)
        return fail(GT_ENOMEM, "the no-allocation entry point needs %zu bytes where its scratch share holds %zu",
                    bytes, b.cap);
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (gt::alloc_should_fail()) return fail(GT_ENOMEM, "device allocation of %zu bytes failed (injected)", bytes);
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(GT_ENOMEM, "device allocation of %zu bytes failed", bytes);
    }
    b.cap = bytes;
    return GT_OK;
}

// As ensure(), and a newly allocated buffer is zeroed on `st` (never inside
// a graph capture: ensure() runs before the pipeline is enqueued).
int ensure_zeroed(DevBuf &b, size_t bytes, hipStream_t st) {
    const void *old = b.p;
    const size_t old_cap = b.cap;
    GT_TRY(ensure(b, bytes));
    if (b.p != old || b.cap != old_cap) GT_HIP(hipMemsetAsync(b.p, 0, b.cap, st));
    return GT_OK;
}

void free_buf(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

struct Tune {
    float lf = 0.35f;
    int dedup = 0;      // 0 auto (LDS tiles when d <= 10 and vd <= 4), 1 on, -1 off
    int tile_w = 0;     // 0 auto: 32, or 16 from d = 7
    int wpc = 0;        // waves-per-CU cap of the tiled and neighbour kernels (0: none)
    bool graph = true;  // replay the pipeline as a cached hipGraph
    bool grid = false;  // the bilateral grid instead of the lattice (GRID entry points only)
    int nbr = 0;        // 0 auto (hash lookups for d >= 11 or few expected vertices), 1 sort + merge, 2 hash
    bool parts = true;  // row-band table parts where they apply
    // a caller-bounded arena (the no-allocation entry point): this table
    // capacity and at most this many vertices, no regrowth
    uint32_t fixed_cap = 0;
    int vert_limit = 0;
};

Tune resolve(const gt_tuning *t) {
    Tune r;
    if (!t) return r;
    if (t->load_factor > 0) r.lf = std::min(0.7f, std::max(0.05f, t->load_factor));
    r.dedup = t->dedup;
    if (t->tile_w != 0) r.tile_w = t->tile_w;
    if (t->waves_per_cu > 0) r.wpc = std::min(64, std::max(12, t->waves_per_cu));
    if (t->graph < 0) r.graph = false;
    if (t->neighbours == 1 || t->neighbours == 2) r.nbr = t->neighbours;
    if (t->parts < 0) r.parts = false;
    return r;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) { (void)hipGetLastError(); return 0.f; }
    return ms;
}

// Per-channel bounds of the positions on the device (one host round trip),
// then the key packing plan.  (The values' bounds, behind the fixed-point
// scales, are taken on the device in every call.)
int make_plan(gt_ctx *c, const gt::PixSrc &src, int n, int d, int vd, bool bilat) {
    const int nplanes = bilat ? src.cr : d;
    const float *planes = bilat ? src.ref : src.pos;
    const int64_t cs = bilat ? src.ref_cs : src.pos_cs;
    const int nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)n + 4095) / 4096));
    GT_TRY(ensure(c->mm_part, (size_t)nblk * std::max(nplanes, 1) * 3 * 4));
    GT_TRY(ensure(c->mm_out, (size_t)std::max(nplanes, 1) * 3 * 4));
    float *part = (float *)c->mm_part.p, *mout = (float *)c->mm_out.p;
    GT_HIP(gt::launch_minmax(planes, cs, nplanes, nullptr, 0, 0, n, part, nblk, mout, c->stream));
    if (nplanes > 0)
        GT_HIP(hipMemcpyAsync(c->h_mm, mout, (size_t)nplanes * 3 * 4, hipMemcpyDeviceToHost, c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    const float *mm = c->h_mm;
    float lo[gt::kMaxDH], hi[gt::kMaxDH];
    for (int i = 0; i < d; i++) {
        float a, b;
        if (i < nplanes) {
            const float sc = bilat ? src.ic : src.inv[i];
            a = mm[3 * i] * sc;
            b = mm[3 * i + 1] * sc;
            if (mm[3 * i + 2] != 0.0f || !(mm[3 * i] <= mm[3 * i + 1])) a = b = INFINITY;  // no usable bound
        } else {
            // spatial channel k of the bilateral form: float(x) * (1/sigma), x in [0, extent-1]
            const int k = i - src.cr;
            int ext[3], m = 0;
            float inv[3];
            if (src.fx) { ext[m] = src.W; inv[m++] = src.ix; }
            if (src.fy) { ext[m] = src.H; inv[m++] = src.iy; }
            if (src.ft) { ext[m] = n / (src.W * src.H); inv[m++] = src.it; }
            a = 0.0f * inv[k];
            b = (float)(ext[k] - 1) * inv[k];
        }
        lo[i] = std::min(a, b);
        hi[i] = std::max(a, b);
    }
    float sf[gt::kMaxDH];
    gt::plan_scale_factors(d, sf);
    gt::plan_from_bounds(d, sf, lo, hi, c->plan);
    c->plan_valid = true;
    c->plan_d = d;
    c->plan_vd = vd;
    c->plan_kind = bilat ? 1 : 0;
    c->plan_n = n;
    c->last_lo = lo[d - 1];
    c->last_hi = hi[d - 1];
    return GT_OK;
}

struct SliceSpec {
    bool enabled = false;
    gt::PixSrc src{};
    int n = 0;
    gt::OutDst dst{};
};

// A batch of bilateral images (gt_joint_bilateral_batch): device descriptors
// and the tile count; the PixSrc/OutDst templates carry the shared fields.
struct BatchSpec {
    gt::BatchArgs ba{};
    int ntiles = 0;
    int Cr = 0, C = 0;
    gt::PixSrc tmpl{};
    gt::OutDst tmpl_dst{};
    uint64_t sig = 0;      // the images' sizes and filters
    bool parts = false;    // the descriptors hold part slices (vertex counts known)
};

template <class T>
void zero(T &x) { memset(&x, 0, sizeof(T)); }

// Appends raw bytes to a graph signature.
template <class T>
void sig_add(std::string &k, const T &v) { k.append((const char *)&v, sizeof(T)); }

// The whole pipeline: splat, compaction, sort, finalize, neighbour table,
// blur, and (optionally) slice, enqueued on the context's stream.
struct Launch {
    int d, vd, storage, W, pw, m_max, vs, n, row_w, tile_w, wpc;
    int m_grid;   // vertices the hash-neighbour and blur grids are sized for (those kernels loop past it)
    bool packed, bilat, dedup, count, hashnbr;
    bool clear_sums;   // clear the fixed-point sums too (not known to be zero)
    bool clear_marks;  // clear the non-finite marks too (a call that stopped early may have left some)
    bool nbr_per_vertex;   // hash neighbours one thread per vertex (else per vertex and axis)
    bool nbr_half;         // ... looking up one sign and scattering the other (entries set to -1 by a fill)
    bool blur_pairs;       // blur passes two per launch
    int nbr_fill_grid;     // one-end form: 0, the finalize fills the entries; > 0, a fill kernel's grid
    uint32_t cap;
    gt::LatConst lc;
    gt::KeyPlan kp;
    gt::ValScale *vsc;   // device: this call's fixed-point scales (computed by the pipeline)
    float *vs_part;      // device: per-workgroup value bounds behind them
    int vs_nblk;
    gt::TableArgs ta;
    gt::PartMap pm;
    float *splat_copy;
    gt::Replay rp;   // rp.rec != nullptr: the splat records, the slice replays
};

// Stage boundary event (eager runs only: a captured event record is a graph
// dependency, not a timestamp, so profiled calls are never captured).
hipError_t mark(gt_ctx *c, int i, hipStream_t s) { return hipEventRecord(c->ev[i], s); }

int enqueue(gt_ctx *c, const Launch &L, const gt::PixSrc &src, const SliceSpec &sl, const BatchSpec *batch,
            int *passes, void **vals_out) {
    hipStream_t s = c->stream;
    gt::Counters *ctr = (gt::Counters *)c->ctr.p;
    const gt::TableArgs &ta = L.ta;
    // this call's value bounds, then the table cleared and the bounds turned
    // into fixed-point scales, on the device
    const gt::ClearArgs cl{ta.tkeys, (size_t)L.cap * L.W, ta.tacc, L.clear_sums ? (size_t)L.cap * L.vd : 0, ta.tnf,
                           L.clear_marks ? (size_t)L.cap : 0, ta.tags, ta.tags ? (size_t)L.cap : 0};
    if (batch)
        GT_HIP(gt::launch_value_scale_batch(cl, batch->ba, batch->ntiles, L.tile_w, batch->C, L.n, L.vd, L.vs_part,
                                            L.vs_nblk, L.vsc, ctr, s));
    else
        GT_HIP(gt::launch_value_scale(cl, src.val, src.val_cs, src.nval, src.homog, L.n, L.vd, L.vs_part, L.vs_nblk,
                                      L.vsc, ctr, s));
    if (c->profiling) GT_HIP(mark(c, 0, s));
    if (batch)
        GT_HIP(gt::launch_splat_batch(L.d, batch->tmpl, batch->ba, batch->ntiles, L.tile_w, L.vd, L.lc, L.kp, L.vsc,
                                      ta, L.rp, ctr, L.count, L.wpc, s));
    else
        GT_HIP(gt::launch_splat(L.d, L.packed, L.bilat, L.dedup, src, L.n, L.row_w, L.tile_w, L.vd, L.lc, L.kp, L.vsc,
                                ta, L.rp, ctr, L.count, L.wpc, s));
    if (c->profiling) GT_HIP(mark(c, 1, s));
    *passes = 0;
    if (L.hashnbr) {   // ids in compaction order: the finalize rides along with the compaction
        GT_HIP(gt::launch_compact_final(L.W, L.storage, ta, L.cap, (uint32_t *)c->counts.p,
                                        (uint32_t *)c->counts.p + chunks_of(L.cap), L.vsc, L.vd, L.vs,
                                        (uint64_t *)c->vkeys.p, c->vals_a.p, ctr, L.m_max,
                                        gt::NbrInit{L.nbr_half ? (int2 *)c->nbr.p : nullptr, L.m_max, L.d + 1,
                                                    L.nbr_fill_grid},
                                        s));
    } else {
        GT_HIP(gt::launch_compact(L.W, ta, L.cap, (uint32_t *)c->counts.p, (uint32_t *)c->slot_list.p,
                                  (uint64_t *)c->pk_a.p, L.packed, ctr, L.m_max, s));
        if (!L.packed)
            GT_HIP(gt::launch_pack(L.d, ta, (const uint32_t *)c->slot_list.p, ctr, L.m_max, L.kp, L.pw,
                                   (uint64_t *)c->pk_a.p, s));
        gt::SortBufs sb{(uint64_t *)c->pk_a.p,         (uint64_t *)c->pk_b.p,    (uint32_t *)c->slot_list.p,
                        (uint32_t *)c->pay_b.p,        (uint32_t *)c->hist.p,    (uint64_t *)c->sort_status.p,
                        (uint32_t *)c->sort_hist.p, kLookbackTiles};
        bool in_b = false;
        GT_HIP(gt::radix_sort(sb, ctr, L.m_max, L.pw, L.kp.bits, &in_b, passes, s));
        const uint32_t *sorted = in_b ? sb.pay_b : sb.pay_a;
        const uint64_t *sorted_keys = L.packed ? (in_b ? sb.pk_b : sb.pk_a) : nullptr;
        GT_HIP(gt::launch_finalize(L.W, L.storage, ta, L.vsc, sorted, sorted_keys, ctr, L.m_max, L.vd, L.vs,
                                   (uint64_t *)c->vkeys.p, c->vals_a.p, L.splat_copy, s));
    }
    if (c->profiling) GT_HIP(mark(c, 2, s));
    int2 *nbr = (int2 *)c->nbr.p;
    if (L.hashnbr)
        GT_HIP(gt::launch_neighbors_hash(L.d, L.packed, L.kp, (const uint64_t *)c->vkeys.p, ctr, L.m_max, L.m_grid,
                                         ta, L.pm, L.nbr_per_vertex, L.nbr_half, nbr, s));
    else
        GT_HIP(gt::launch_neighbors(L.d, L.packed, L.kp, (const uint64_t *)c->vkeys.p, ctr, L.m_max, nbr, L.wpc,
                                    s));
    if (c->profiling) GT_HIP(mark(c, 3, s));
    void *cur = c->vals_a.p, *nxt = c->vals_b.p;
    // (the blur passes loop over their vertices too; a grid for the expected
    // count leaves fewer workgroups with nothing to do: 64-image batch 16.43
    // -> 16.29 ms)
    const int blur_grid = std::max(1, L.m_grid / 2);
    for (int j = 0; j <= L.d;) {
        if (L.blur_pairs && j + 1 <= L.d) {
            GT_HIP(gt::launch_blur2(L.vs, nbr + (size_t)j * L.m_max, nbr + (size_t)(j + 1) * L.m_max, ctr, blur_grid,
                                    L.pm, (const float *)cur, (float *)nxt, s));
            j += 2;
        } else {
            GT_HIP(gt::launch_blur(L.storage, L.vs, j, L.vsc, nbr + (size_t)j * L.m_max, ctr, blur_grid, L.pm, cur,
                                   nxt, s));
            j++;
        }
        std::swap(cur, nxt);
    }
    if (c->profiling) GT_HIP(mark(c, 4, s));
    const unsigned *sbits = L.storage ? &L.vsc->vmax_bits[L.d][0] : nullptr;   // the last pass's scale
    if (L.rp.rec && batch)
        GT_HIP(gt::launch_slice_replay(L.d, true, L.storage, batch->tmpl, 0, 0, L.tile_w, batch->ba, batch->ntiles,
                                       L.vd, L.lc, L.kp, ta, L.rp, cur, sbits, batch->tmpl_dst, ctr, c->h_ctr,
                                       L.wpc, s));
    else if (L.rp.rec)
        GT_HIP(gt::launch_slice_replay(L.d, sl.src.ref != nullptr, L.storage, sl.src, sl.n, L.row_w, L.tile_w,
                                       gt::BatchArgs{nullptr, 0}, 0, L.vd, L.lc, L.kp, ta, L.rp, cur, sbits, sl.dst,
                                       ctr, c->h_ctr, L.wpc, s));
    else if (batch)
        GT_HIP(gt::launch_slice_batch(L.d, L.storage, batch->tmpl, batch->ba, batch->ntiles, L.tile_w, L.vd, L.lc,
                                      L.kp, ta, cur, sbits, batch->tmpl_dst, ctr, L.count, L.wpc, s));
    else if (sl.enabled)
        GT_HIP(gt::launch_slice(L.d, L.packed, sl.src.ref != nullptr, L.dedup, L.storage, sl.src, sl.n, L.row_w,
                                L.tile_w, L.vd, L.vs, L.lc, L.kp, ta, cur, sbits, sl.dst, ctr, L.count, L.wpc, s));
    if (c->profiling) GT_HIP(mark(c, 5, s));
    if (!L.rp.rec)   // (the replay slice writes them itself)
        GT_HIP(hipMemcpyAsync(c->h_ctr, ctr, sizeof(gt::Counters), hipMemcpyDeviceToHost, s));
    *vals_out = cur;
    return GT_OK;
}

void drop_graph(gt_ctx *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gsig.clear();
}

// Runs the enqueue sequence, through a cached hipGraph when `graph` is set:
// the graph is captured once per signature (every value baked into a launch)
// and replayed afterwards, one launch for the whole pipeline.
int execute(gt_ctx *c, const Launch &L, const gt::PixSrc &src, const SliceSpec &sl, const BatchSpec *batch,
            bool graph, int *passes, void **vals_out) {
    if (!graph) return enqueue(c, L, src, sl, batch, passes, vals_out);
    std::string sig;
    sig_add(sig, L);
    sig_add(sig, src);
    sig_add(sig, sl.enabled);
    sig_add(sig, sl.src);
    sig_add(sig, sl.n);
    sig_add(sig, sl.dst);
    sig_add(sig, c->profiling);
    sig_add(sig, c->stream);
    const void *bufs[] = {c->tkeys.p,  c->tacc.p,  c->tnf.p, c->vscale.p, c->vs_part.p, c->ctr.p,         c->counts.p,    c->slot_list.p,
                          c->slot_id.p, c->pk_a.p,       c->pk_b.p,        c->pay_b.p,     c->hist.p,
                          c->sort_status.p, c->sort_hist.p, c->vkeys.p,    c->vals_a.p,    c->vals_b.p,
                          c->nbr.p};
    sig_add(sig, bufs);
    if (batch) {
        sig_add(sig, batch->ba);
        sig_add(sig, batch->ntiles);
        sig_add(sig, batch->tmpl);
        sig_add(sig, batch->tmpl_dst);
    }
    if (!(c->gexec && c->gsig == sig)) {
        drop_graph(c);
        hipGraph_t g = nullptr;
        GT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(c, L, src, sl, batch, &c->gpasses, &c->gvals);
        hipGraph_t captured = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &captured);
        if (rc != GT_OK) {
            if (captured) (void)hipGraphDestroy(captured);
            return rc;
        }
        if (e != hipSuccess) return fail(GT_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
        g = captured;
        const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->gexec = nullptr;
            return fail(GT_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
        }
        c->gsig = sig;
    }
    GT_HIP(hipGraphLaunch(c->gexec, c->stream));
    *passes = c->gpasses;
    *vals_out = c->gvals;
    return GT_OK;
}

int make_plan_batch(gt_ctx *c, const BatchSpec &b, const std::vector<gt::BatchImage> &imgs, int d, int vd,
                    int tile_w);

template <int N>
float density_find(const gt_density (&e)[N], uint64_t key) {
    for (const gt_density &x : e)
        if (key && x.key == key) return x.per_pixel;
    return 0.f;
}

template <int N>
void density_store(gt_density (&e)[N], int &next, uint64_t key, float per_pixel) {
    gt_density *slot = nullptr;
    for (gt_density &x : e)
        if (x.key == key) slot = &x;
    if (!slot) {
        slot = &e[next];
        next = (next + 1) % N;
    }
    slot->key = key;
    slot->per_pixel = per_pixel;
}

float density_get(gt_ctx *c, uint64_t key) {
    if (!key) return 0.f;
    if (gt_density_table *t = c->shared_density) {
        std::lock_guard<std::mutex> lk(t->mu);
        return density_find(t->e, key);
    }
    return density_find(c->density, key);
}

void density_put(gt_ctx *c, uint64_t key, float per_pixel) {
    if (gt_density_table *t = c->shared_density) {
        std::lock_guard<std::mutex> lk(t->mu);
        density_store(t->e, t->next, key, per_pixel);
    } else {
        density_store(c->density, c->density_next, key, per_pixel);
    }
}

// Signature of a filter for the vertex-density memory: the position form,
// d, vd and the scales of every position channel (FNV-1a over their bytes;
// never 0).
uint64_t density_key(const gt::PixSrc &src, int kind, int d, int vd) {
    std::string k;
    sig_add(k, kind);
    sig_add(k, d);
    sig_add(k, vd);
    if (src.ref) {
        sig_add(k, src.cr);
        sig_add(k, src.ic);
        const float spatial[3] = {src.fx ? src.ix : 0.f, src.fy ? src.iy : 0.f, src.ft ? src.it : 0.f};
        sig_add(k, spatial);
    } else {
        for (int i = 0; i < d; i++) sig_add(k, src.inv[i]);
    }
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : k) h = (h ^ ch) * 1099511628211ull;
    return h ? h : 1;
}

int run(gt_ctx *c, const gt::PixSrc &src, int n, int row_w, int d, int vd, int storage, const Tune &tu,
        const SliceSpec &sl, bool keep_splat, gt_stats *st, const BatchSpec *batch = nullptr,
        const std::vector<gt::BatchImage> *bimgs = nullptr) {
    hipStream_t s = c->stream;
    c->have = false;   // until this call completes (a failed call may have freed the arrays)
    const bool bilat = batch ? true : src.ref != nullptr;
    const bool count = c->profiling >= 2;
    const bool dedup = batch ? true : (tu.dedup == 0) ? (d <= 10 && vd <= 4) : (tu.dedup > 0);
    const int kind = batch ? 2 : bilat ? 1 : 0;

    // Expected vertex count: the vertices per pixel this filter gave before
    // (any image size), else the last call's M for the same size and d, else
    // the pixel count (a guess: M / n runs from ~0.05 to d + 1).
    const uint64_t dkey = batch ? 0 : density_key(src, kind, d, vd);
    const float per_pixel = density_get(c, dkey);
    uint64_t est;
    if (per_pixel > 0.f) est = std::max<uint64_t>(65536, (uint64_t)((double)n * per_pixel * 1.05) + 64);
    else if (c->last_n == n && c->last_d == d && c->last_M > 0) est = (uint64_t)(c->last_M * 1.05) + 64;
    else est = std::max<uint64_t>(65536, (uint64_t)n);
    est = std::min<uint64_t>(est, (uint64_t)n * (d + 1));
    uint32_t cap = next_pow2((uint64_t)(est / tu.lf));
    uint32_t cap_limit = next_pow2((uint64_t)((double)n * (d + 1) / tu.lf));
    // Tile width: the caller's, else 32 pixels, or 16 from d = 7, where a
    // tile's pixels reach more distinct vertices (non-local means, d = 8:
    // 2.237 -> 2.201 ms; the 2048^2 joint bilateral keeps 32: 0.377 against
    // 0.380 ms).
    const int tile_w = tu.tile_w ? tu.tile_w : (d >= 7 ? 16 : 32);
    if (cap > cap_limit) cap = cap_limit;
    if (tu.fixed_cap) cap = cap_limit = tu.fixed_cap;
    auto replan = [&]() -> int {
        if (batch) return make_plan_batch(c, *batch, *bimgs, d, vd, tile_w);
        return make_plan(c, src, n, d, vd, bilat);
    };
    const int images = batch ? batch->ba.count : 1;
    // A Gauss transform over an image (row_w > 0) may split its table by its
    // last position channel: that needs the channel's bounds for this call.
    const bool gauss_rows = tu.parts && !batch && !bilat && row_w > 0 && n >= (1 << 20);
    if (!(c->plan_valid && c->plan_d == d && c->plan_vd == vd && c->plan_kind == kind &&
          (images <= 1 || (images - 1) >> c->plan.img_bits == 0) && (!gauss_rows || c->plan_n == n)))
        GT_TRY(replan());
    bool parts_failed = false;   // a Gauss transform's split table filled up: redo on the plain one

    int retries = 0, passes = 0;
    for (;;) {
        Launch L;
        zero(L);
        L.d = d;
        L.vd = vd;
        L.storage = storage;
        L.n = n;
        L.row_w = row_w;
        L.tile_w = tile_w;
        L.wpc = tu.wpc;
        L.bilat = bilat;
        L.dedup = dedup;
        L.count = count;
        L.cap = cap;
        gt::plan_scale_factors(d, L.lc.sf);
        L.kp = c->plan;
        L.packed = L.kp.packed != 0;
        if (batch && !L.packed) return fail(GT_EINVAL, "batched lattice keys do not fit 63 bits");
        L.W = L.packed ? 1 : (d + 3) / 4;
        // Row bands: in a one-frame bilateral form the last position channel
        // is the row, y / sigma_s, and the last lattice coordinate of every
        // vertex is -d sf y / sigma_s to within 2(d + 1).  P equal ranges of
        // it give P table parts (P a multiple of 8); the splat and slice hand
        // XCD x the x-th eighth of the tiles, in row order, and (ids in
        // compaction order) the neighbour and blur kernels the ids of parts
        // [x P / 8, (x + 1) P / 8), in order, so the table lines, sums and
        // values each XCD touches at a time stay mostly in its own L2.
        // A Gauss transform over an image is split the same way by its last
        // position channel when the caller laid that channel out by rows (an
        // uneven split only lengthens probes; one that fills the table is
        // redone without parts).
        bool rows = false;
        double plo = 0, phi = 0;   // bounds of the last position channel, scaled
        if (tu.parts && !parts_failed && !batch && cap >= (1u << 14)) {
            if (bilat && src.fy && !src.ft) {
                rows = true;
                phi = (double)(src.H - 1) * src.iy;
            } else if (gauss_rows && std::isfinite(c->last_lo) && std::isfinite(c->last_hi) &&
                       c->last_hi > c->last_lo) {
                rows = true;
                plo = c->last_lo;
                phi = c->last_hi;
            }
        }
        if (rows) {
            // fields hold c_i - lo_i (packed) or c_i + 32768 (wide)
            double sum_lo = 0;
            for (int i = 0; i < d; i++) sum_lo += L.packed ? L.kp.lo[i] : -32768;
            const int margin = 2 * (d + 1);
            const double base = std::floor((double)d * L.lc.sf[d - 1] * plo);
            const double range = std::ceil((double)d * L.lc.sf[d - 1] * phi - base) + 2.0 * margin;
            // About 2 MB of key words per part (each XCD's L2 is 4 MB), and
            // parts at least 4 units of c_d wide; at least eight.  (A vertex's
            // neighbours are within d units of its c_d: parts narrower than
            // that still hold them a few parts, so a few ids, apart.)
            int lg = 0;
            while ((1u << lg) < cap) lg++;
            int plg = 3;
            while (plg < lg - 8 && ((uint64_t)cap * L.W * 8 >> (plg + 1)) >= (1u << 21) &&
                   range / (double)(1 << (plg + 1)) >= 4.0)
                plg++;
            if (range >= (double)(1 << plg) && range < (double)(1 << 24)) {
                L.ta.part_base = (int)(sum_lo - base) + margin;
                L.ta.part_range = (int)range;
                L.ta.part_mul = (uint32_t)((double)(1ull << (24 + plg)) / (range + 1.0));
                L.ta.part_lg = plg;
                L.ta.part_shift = lg - plg;
            }
        }
        // A batch whose images' vertex counts are known: each image takes a
        // run of parts in proportion to its vertices, split by its rows (the
        // descriptors hold the slices; the key's tag picks the image).
        if (tu.parts && !parts_failed && batch && batch->parts && L.packed && cap >= (1u << 14)) {
            double sum_lo = 0;
            for (int i = 0; i < d; i++) sum_lo += L.kp.lo[i];
            int lg = 0;
            while ((1u << lg) < cap) lg++;
            int plg = 3;
            while (plg < lg - 8 && ((uint64_t)cap * L.W * 8 >> (plg + 1)) >= (1u << 21)) plg++;
            L.ta.part_img = batch->ba.img;
            L.ta.part_base = (int)sum_lo + 2 * (d + 1);
            L.ta.part_mul = 1;   // on; the per-image descriptors carry the scales
            L.ta.part_lg = plg;
            L.ta.part_shift = lg - plg;
        }
        // Vertex ids: sorted (merge neighbours) or compaction order (hash
        // neighbours, no sort); the output is the same bit for bit.  With
        // table parts the hash form wins at every size measured (2048^2
        // joint bilateral, 0.41 M vertices: 0.45 against 0.65 ms; d = 8,
        // 2.1 M: 2.55 against 3.12 ms; 4096^2 RGB at sigma_s 4, 28.9 M: 16.4
        // against 17.2 ms).  Without them it pays for small lattices (the
        // sort's ~30 launches dominate) and at high d (many passes over wide
        // keys: d = 16, 17.8 M vertices, 37.4 -> 31.3 ms); in between, sorted
        // ids keep the blur's gathers local.  A kept lattice
        // (gt_build_lattice) is always sorted.
        L.hashnbr = !keep_splat && (tu.nbr == 2 || (tu.nbr == 0 && (L.ta.part_mul != 0 ||
                                                                   (!batch && (d >= 11 || est <= (1u << 20))))));
        // One thread per vertex shares the key work of its d + 1 axes, one per
        // (vertex, axis) has d + 1 times the threads: the latter for small,
        // low-d lattices (512^2 bilateral, 0.17 M vertices: 29 against
        // 35 us; 2048^2 joint bilateral, 0.41 M: 57 against 55 us; 4096^2
        // RGB at sigma_s 4, 28.9 M: 5.6 against 4.3 ms).
        L.nbr_per_vertex = d >= 7 || est >= (1u << 18);
        // Each lattice edge looked up from one end only: u = k + 1 - (d+1)e_j
        // found from k gives both u (k's first entry) and k (u's second),
        // for a -1 fill of every entry and a scattered store.  That pays
        // where lookups are dear (d >= 7: wide keys; d = 16 transform 16.8
        // -> 14.8 ms, non-local means 2.30 -> 2.23 ms) and where the
        // scattered stores stay in L2 (2048^2 joint bilateral: neighbours
        // 52 -> 41 us plus a 7 us fill); not for large d = 5 lattices (a
        // 64-image batch, 31 M vertices: 16.4 -> 17.6 ms).
        L.nbr_half = L.hashnbr && L.nbr_per_vertex && (d >= 7 || est * (d + 1) * 8 <= (32u << 20));
        // Two blur passes per launch (FP32 values only: a narrow storage
        // rounds between passes) where the lattice is sparse, more vertices
        // than (d+1)/2 per pixel, so that most neighbours are missing and the
        // nine reads of a fused pair are few: d = 16 transform 14.8 -> 14.3
        // ms.  Dense lattices read more than the two launches do (2048^2
        // joint bilateral 0.376 -> 0.384 ms, 64-image batch 16.4 -> 17.0 ms).
        L.blur_pairs = storage == 0 && vd <= 4 && est * 2 > (uint64_t)n * (d + 1);
        L.m_max = (int)std::min<uint64_t>(cap, (uint64_t)n * (d + 1));
        if (tu.vert_limit > 0) L.m_max = std::min(L.m_max, tu.vert_limit);
        // Grids for twice the expected count, not the table's capacity (a
        // batch's 2^27 slots for 31 M vertices start 4x the workgroups
        // needed, in every blur pass); the margin keeps the table part with
        // the most vertices within one pass of its workgroups.
        L.m_grid = (int)std::min<uint64_t>((uint64_t)L.m_max, 2 * est);
        // The one-end form's -1 fill: by the finalize's workgroups, each its
        // chunk's ids, while the entries are few (2048^2 joint bilateral,
        // 20 MB: 3 us there against a 7 us launch); by a kernel of its own,
        // coalesced over whole rows, for large tables (d = 16, 2.4 GB: 0.3 ms
        // less).
        L.nbr_fill_grid = est * (d + 1) * 8 <= (64u << 20) ? 0 : std::max(L.m_grid, 1);
        L.pw = L.packed ? 1 : std::max(1, (L.kp.bits + 63) / 64);
        L.vs = value_stride(vd);
        const unsigned ncount = (cap + 2047) / 2048;
        GT_TRY(ensure(c->tkeys, (size_t)cap * L.W * 8));
        const void *old_acc = c->tacc.p;
        GT_TRY(ensure(c->tacc, (size_t)cap * vd * 8));
        if (c->tacc.p != old_acc) c->acc_zero = 0;
        // The sums need clearing unless a complete earlier call left this
        // much of the array zero; until this call completes, none is known.
        const size_t acc_bytes = (size_t)cap * vd * 8;
        const size_t acc_zero = c->acc_zero;
        L.clear_sums = acc_zero < acc_bytes;
        c->acc_zero = 0;
        // non-finite marks: zero unless an earlier call stopped before its
        // finalize read them (a full table); until this call completes, unknown
        {
            const void *old_tnf = c->tnf.p;
            GT_TRY(ensure_zeroed(c->tnf, (size_t)cap * 4, s));
            if (c->tnf.p != old_tnf) c->tnf_clean = true;
        }
        L.clear_marks = !c->tnf_clean;
        c->tnf_clean = false;
        L.vs_nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)n + 4095) / 4096));
        GT_TRY(ensure(c->vs_part, (size_t)L.vs_nblk * std::max(vd, 1) * 3 * 4));
        GT_TRY(ensure(c->vscale, sizeof(gt::ValScale)));
        L.vsc = (gt::ValScale *)c->vscale.p;
        L.vs_part = (float *)c->vs_part.p;
        GT_TRY(ensure(c->ctr, sizeof(gt::Counters)));
        GT_TRY(ensure(c->counts, (size_t)ncount * 8));   // counts, then offsets (a self-scanning finalize)
        GT_TRY(ensure(c->slot_id, (size_t)cap * 4));
        if (!L.hashnbr) {   // the slot list and the radix sort's buffers
            GT_TRY(ensure(c->slot_list, (size_t)L.m_max * 4));
            GT_TRY(ensure(c->pk_a, (size_t)L.pw * L.m_max * 8));
            GT_TRY(ensure(c->pk_b, (size_t)L.pw * L.m_max * 8));
            GT_TRY(ensure(c->pay_b, (size_t)L.m_max * 4));
            GT_TRY(ensure_zeroed(c->hist, gt::sort_state_words() * 4, s));
            GT_TRY(ensure_zeroed(c->sort_status, gt::sort_status_entries(L.m_max, L.pw) * 8, s));
            GT_TRY(ensure(c->sort_hist, gt::sort_hist_words(L.m_max, L.pw) * 4));
        }
        GT_TRY(ensure(c->vkeys, (size_t)L.m_max * L.W * 8));
        // vals_a takes the finalize's FP32 rows (every storage), vals_b the
        // storage type; the passes alternate between them
        GT_TRY(ensure(c->vals_a, (size_t)L.m_max * L.vs * 4));
        GT_TRY(ensure(c->vals_b, (size_t)L.m_max * L.vs * store_bytes(storage)));
        GT_TRY(ensure(c->nbr, (size_t)(d + 1) * L.m_max * 8));
        // Replay slice: a tiled splat records each pixel's tile-table slots,
        // and the slice at the same positions reads them back instead of
        // building the keys and inserting them into its own tile table
        // (2048^2 joint bilateral: slice 116 -> 59 us, splat 146 -> 149 us,
        // 0.429 -> 0.380 ms; recording the weights as well made the splat
        // 175 us for a 52 us slice).  Not in the no-allocation entry point,
        // whose buffers are fixed shares of its scratch array.
        zero(L.rp);
        const bool same_pos = sl.enabled && sl.n == n && sl.src.pos == src.pos && sl.src.ref == src.ref;
        if (!g_no_alloc && L.dedup && vd <= 4 && d <= 10 && (batch || same_pos)) {
            int tw = L.tile_w;
            const int64_t tiles = batch ? batch->ntiles : gt::pixel_tiles(n, row_w, tw);
            const int64_t stride = tiles * gt::kTilePixels;
            GT_TRY(ensure(c->rp_rec, (size_t)gt::replay_dwords(d) * stride * 4));
            GT_TRY(ensure(c->rp_gs, (size_t)(d + 1) * stride * 4));
            GT_TRY(ensure(c->rp_tent, (size_t)tiles * gt::kTileKeys * 8));
            GT_TRY(ensure(c->rp_tcnt, (size_t)tiles * 4));
            L.rp = {(uint32_t *)c->rp_rec.p, (uint32_t *)c->rp_gs.p, (uint2 *)c->rp_tent.p, (uint32_t *)c->rp_tcnt.p,
                    stride};
        }
        L.splat_copy = nullptr;
        if (keep_splat) {
            GT_TRY(ensure(c->splat_copy, (size_t)L.m_max * vd * 4));
            L.splat_copy = (float *)c->splat_copy.p;
        }
        L.ta.tkeys = (uint64_t *)c->tkeys.p;
        L.ta.tacc = (unsigned long long *)c->tacc.p;
        L.ta.tnf = (uint32_t *)c->tnf.p;
        L.ta.mask = cap - 1;
        L.ta.slot_id = (int *)c->slot_id.p;
        // Tags for the neighbour lookups of wide keys (hash ids): d = 16,
        // 1024^2 (17.8 M vertices, 2 GB of key words; 94% of lookups miss):
        // neighbours 12.1 -> 8.1 ms probing tag by tag, 6.5 ms eight tags
        // at a time.
        L.ta.tags = nullptr;
        if (L.hashnbr && L.W > 1) {
            GT_TRY(ensure(c->tags, (size_t)cap));
            L.ta.tags = (uint8_t *)c->tags.p;
        }
        // One-word keys leave room above the key for the vertex id: fold it in
        // at the finalize, so lookups after it read key and id in one word.
        {
            int idbits = 1;
            while (idbits < 31 && (1u << idbits) < (uint32_t)L.m_max) idbits++;
            const bool fold = L.packed && L.kp.bits + idbits <= 63;
            L.ta.idshift = fold ? L.kp.bits : 0;
            L.ta.keymask = fold ? (1ull << L.kp.bits) - 1 : ~0ull;
        }
        if (L.hashnbr && L.ta.part_mul)
            L.pm = {(const uint32_t *)c->counts.p + (ncount <= (unsigned)gt::kSelfScanChunks ? ncount : 0u),
                    (int)(cap / 8 / gt::kCompactSlots)};

        void *cur = nullptr;
        GT_TRY(execute(c, L, src, sl, batch, tu.graph && !keep_splat && c->profiling == 0, &passes, &cur));
        GT_HIP(hipStreamSynchronize(s));
        const gt::Counters hc = *c->h_ctr;
        // every mark was read by the finalize unless the table filled up
        if (!((hc.flags & gt::kFlagFull) && (hc.flags & gt::kFlagNonFinite))) c->tnf_clean = true;
        const unsigned hard = hc.flags & ~gt::kFlagNonFinite;

        if (hc.flags & gt::kFlagBad)
            return fail(GT_ERANGE, "position is NaN/Inf or a lattice key leaves the int16 range");
        if (hc.flags & gt::kFlagSort) return fail(GT_EDEVICE, "radix sort look-back timed out");
        if (hc.flags & gt::kFlagWait) return fail(GT_EDEVICE, "a hash table key word never appeared");
        if (hc.flags & gt::kFlagRange) {
            // Keys outside the cached plan: bound these positions again.
            if (++retries > 6) return fail(GT_EDEVICE, "inputs keep leaving the computed bounds");
            GT_TRY(replan());
            continue;
        }
        if (hc.flags & gt::kFlagFull) {
            if (++retries > 6) return fail(GT_ENOMEM, "lattice does not fit a hash table of %u slots", cap);
            // a batch's measured vertex counts no longer fit: the sorted redo
            // measures them again
            if (L.ta.part_img) c->batch_sig = 0;
            // One part of a split table filled while the table as a whole held
            // room (an uneven split: a skewed last channel, rows of uneven
            // content, stale batch counts): redo at this capacity without
            // parts.  A table that is really too small fills to ~0.9 before an
            // insert runs out of probes.
            if (L.ta.part_mul && !parts_failed && (double)hc.filled < 0.6 * (double)cap) {
                parts_failed = true;
                continue;
            }
            if (tu.fixed_cap)
                return fail(GT_ENOMEM, "lattice exceeds the no-allocation maximum of %d vertices (%u table slots)",
                            L.m_max, cap);
            if (cap >= cap_limit) return fail(GT_ENOMEM, "lattice does not fit a hash table of %u slots", cap);
            // a full table held about `cap` vertices: decide the redo's
            // neighbour form on that
            est = std::max<uint64_t>(est, cap);
            // (a bilateral form's rows hold equal pixel counts, so its parts
            // stay; a Gauss transform's last channel may be skewed, and a
            // batch's images may hold other vertex counts than measured)
            if (L.ta.part_mul && (!bilat || batch)) parts_failed = true;
            // M <= (d + 1) n, and a first table holds n / load_factor slots: x4
            // covers every lattice up to d = 7 at the default load in one redo
            // (a larger table only lengthens the redo's clear and compaction)
            cap = std::min<uint32_t>(cap_limit, cap * (d <= 7 ? 4u : 8u));
            continue;
        }
        const int M = (int)hc.M;
        if (hard == 0) c->acc_zero = std::max(acc_zero, acc_bytes);   // the finalize zeroed what the splat summed
        c->last_n = n;
        c->last_d = d;
        c->last_M = (unsigned)M;
        if (dkey) density_put(c, dkey, (float)M / (float)n);
        if (batch && batch->sig && !L.hashnbr && M > 0 && L.packed && c->batch_sig != batch->sig) {
            // sorted ids: the images' vertices are consecutive runs, by tag
            const int nb = batch->ba.count;
            GT_TRY(ensure(c->tag_starts, (size_t)(nb + 1) * 4));
            GT_HIP(gt::launch_tag_starts((const uint64_t *)c->vkeys.p, M, L.kp.coord_bits, nb,
                                         (uint32_t *)c->tag_starts.p, s));
            std::vector<uint32_t> starts(nb + 1);
            GT_HIP(hipMemcpyAsync(starts.data(), c->tag_starts.p, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
            c->batch_verts.assign(nb, 0);
            for (int i = 0; i < nb; i++) c->batch_verts[i] = starts[i + 1] - starts[i];
            c->batch_sig = batch->sig;
        }
        c->have = !L.hashnbr;   // a lattice in compaction order is not exported
        c->M = M;
        c->d = d;
        c->vd = vd;
        c->vs = L.vs;
        c->storage = storage;
        c->packed = L.packed;
        c->cap = cap;
        c->vals_cur = cur;
        if (st) {
            st->M = M;
            st->capacity = (int)std::min<uint32_t>(cap, 0x7fffffff);
            st->retries = retries;
            st->sort_passes = passes;
            if (c->profiling) {
                st->ms_splat = elapsed(c->ev[0], c->ev[1]);
                st->ms_sort = elapsed(c->ev[1], c->ev[2]);
                st->ms_nbr = elapsed(c->ev[2], c->ev[3]);
                st->ms_blur = elapsed(c->ev[3], c->ev[4]);
                st->ms_slice = elapsed(c->ev[4], c->ev[5]);
                st->ms_total = elapsed(c->ev[0], c->ev[5]);
            }
            if (count) {
                st->probes = hc.probes;
                st->inserts = hc.inserts;
                st->atomics = hc.atomics;
            }
        }
        return GT_OK;
    }
}

// Bounds of a batch's guides: one device pass over every tile of every image.
int make_plan_batch(gt_ctx *c, const BatchSpec &b, const std::vector<gt::BatchImage> &imgs, int d, int vd,
                    int tile_w) {
    const int np = b.Cr;
    const int nblk = std::max(1, std::min(1024, b.ntiles));
    GT_TRY(ensure(c->mm_part, (size_t)nblk * std::max(np, 1) * 3 * 4));
    GT_TRY(ensure(c->mm_out, (size_t)std::max(np, 1) * 3 * 4));
    GT_HIP(gt::launch_bounds_batch(b.ba, b.ntiles, tile_w, b.Cr, 0, (float *)c->mm_part.p, nblk,
                                   (float *)c->mm_out.p, c->stream));
    GT_HIP(hipMemcpyAsync(c->h_mm, c->mm_out.p, (size_t)np * 3 * 4, hipMemcpyDeviceToHost, c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    const float *mm = c->h_mm;
    float lo[gt::kMaxDH], hi[gt::kMaxDH];
    for (int i = 0; i < b.Cr; i++) {
        lo[i] = mm[3 * i];
        hi[i] = mm[3 * i + 1];
        if (mm[3 * i + 2] != 0.0f || !(lo[i] <= hi[i])) lo[i] = hi[i] = INFINITY;
    }
    float xmax = 0.0f, ymax = 0.0f;
    for (const gt::BatchImage &bi : imgs) {
        xmax = std::max(xmax, (float)(bi.W - 1) * bi.ix);
        ymax = std::max(ymax, (float)(bi.H - 1) * bi.iy);
    }
    lo[b.Cr] = 0.0f;
    hi[b.Cr] = xmax;
    lo[b.Cr + 1] = 0.0f;
    hi[b.Cr + 1] = ymax;
    float sf[gt::kMaxDH];
    gt::plan_scale_factors(d, sf);
    gt::plan_from_bounds(d, sf, lo, hi, c->plan, (int)imgs.size());
    c->plan_valid = true;
    c->plan_d = d;
    c->plan_vd = vd;
    c->plan_kind = 2;
    return GT_OK;
}

// Sums the statistics of one group of a call run as several lattices.
void add_stats(gt_stats &t, const gt_stats &g) {
    t.M = g.M;
    t.capacity = std::max(t.capacity, g.capacity);
    t.retries += g.retries;
    t.sort_passes += g.sort_passes;
    t.ms_splat += g.ms_splat;
    t.ms_sort += g.ms_sort;
    t.ms_nbr += g.ms_nbr;
    t.ms_blur += g.ms_blur;
    t.ms_slice += g.ms_slice;
    t.ms_total += g.ms_total;
    t.probes += g.probes;
    t.inserts += g.inserts;
    t.atomics += g.atomics;
}

int check_dims(int d, int vd) {
    if (d < 1 || d > GT_MAX_D) return fail(GT_EINVAL, "position dimension d=%d outside [1, %d]", d, GT_MAX_D);
    if (vd < 1 || vd > GT_MAX_VD) return fail(GT_EINVAL, "value dimension vd=%d outside [1, %d]", vd, GT_MAX_VD);
    return GT_OK;
}

int check_storage(gt_storage s) {
    if (s != GT_STORE_F32 && s != GT_STORE_BF16 && s != GT_STORE_F16) return fail(GT_EINVAL, "bad storage %d", (int)s);
    return GT_OK;
}

// The lattice part of JointBilateral on a dense view: positions from the
// guide, values [im, 1], normalised output (src/GaussTransform.cpp:540-677).
// The bilateral grid (GaussTransform::apply, GRID case,
// src/GaussTransform.cpp:227-297): bounds of position * sf over the splat
// (and a distinct slice) positions, grid sizes ceil(max - min) + 1
// (src/DenseGrid.h:159-169), fixed-point scales from the value bounds, then
// zero / splat / finish / blur / slice on the device.  `slice` equal to
// `src` slices at the splat positions.
int run_grid(gt_ctx *c, const gt::PixSrc &src, int n, int d, int vd, const gt::PixSrc &slice, int n_slice,
             const gt::OutDst &dst, bool bilat, gt_stats *st) {
    if (d > gt::kMaxGridD)
        return fail(GT_EINVAL, "the bilateral grid takes at most %d position channels (got %d)", gt::kMaxGridD, d);
    const bool same = !bilat && (slice.pos == src.pos) && n_slice == n;
    const int np = bilat ? src.cr : d;
    const float *planes = bilat ? src.ref : src.pos;
    const int64_t pcs = bilat ? src.ref_cs : src.pos_cs;
    const int nval = src.nval;
    const int nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)std::max(n, n_slice) + 4095) / 4096));
    GT_TRY(ensure(c->mm_part, (size_t)nblk * (np + nval + d) * 3 * 4));
    GT_TRY(ensure(c->mm_out, (size_t)(np + nval + d) * 3 * 4));
    float *part = (float *)c->mm_part.p, *mout = (float *)c->mm_out.p;
    if (n > 0)
        GT_HIP(gt::launch_minmax(planes, pcs, np, src.val, src.val_cs, nval, n, part, nblk, mout, c->stream));
    const bool sep = !bilat && !same && n_slice > 0;
    if (sep)
        GT_HIP(gt::launch_minmax(slice.pos, slice.pos_cs, d, nullptr, 0, 0, n_slice, part, nblk,
                                 mout + (size_t)(np + nval) * 3, c->stream));
    GT_HIP(hipMemcpyAsync(c->h_mm, mout, (size_t)(np + nval + (sep ? d : 0)) * 3 * 4, hipMemcpyDeviceToHost,
                          c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    const float *mm = c->h_mm, *vm = c->h_mm + (size_t)np * 3, *sm = c->h_mm + (size_t)(np + nval) * 3;
    // NaN/Inf values spread as in the reference (per-element marks, see
    // k_grid_splat); the scales come from the finite values
    bool nonfinite = false;
    for (int k = 0; k < (n > 0 ? nval : 0); k++) nonfinite |= vm[3 * k + 2] != 0.0f;
    gt::GridArgs g{};
    g.vd = vd;
    g.sf = sqrtf((float)(1.0 / 3 + (5 - 1) * 0.25));   // taps = 5 (src/GaussTransform.cpp:229)
    float mx[gt::kMaxGridD];
    for (int i = 0; i < d; i++) {
        float lo = INFINITY, hi = -INFINITY;
        auto take = [&](float raw_lo, float raw_hi, float scale, bool bad) -> bool {
            if (bad || !(raw_lo <= raw_hi)) return false;
            const float a = (raw_lo * scale) * g.sf, b = (raw_hi * scale) * g.sf;
            lo = std::min(lo, std::min(a, b));
            hi = std::max(hi, std::max(a, b));
            return true;
        };
        bool ok = true;
        if (n > 0) {
            if (bilat && i >= src.cr) {
                // x * (1/sigma_x) etc. over the image extent
                const int k = i - src.cr;
                int ext[3], m = 0;
                float inv[3];
                if (src.fx) { ext[m] = src.W; inv[m++] = src.ix; }
                if (src.fy) { ext[m] = src.H; inv[m++] = src.iy; }
                if (src.ft) { ext[m] = n / (src.W * src.H); inv[m++] = src.it; }
                ok = take(0.0f, (float)(ext[k] - 1), inv[k], false);
            } else {
                ok = take(mm[3 * i], mm[3 * i + 1], bilat ? src.ic : src.inv[i], mm[3 * i + 2] != 0.0f);
            }
        }
        if (ok && sep) ok = take(sm[3 * i], sm[3 * i + 1], slice.inv[i], sm[3 * i + 2] != 0.0f);
        if (!ok || !std::isfinite(lo) || !std::isfinite(hi))
            return fail(GT_ERANGE, "position channel %d holds NaN or Inf", i);
        g.mn[i] = lo;
        mx[i] = hi;
    }
    int64_t cells = 1;
    for (int i = 0; i < d; i++) {
        const float ext = std::ceil(mx[i] - g.mn[i]) + 1;
        if (!(ext <= 32767.0f))
            return fail(GT_ERANGE, "grid dimension %d spans %g cells (more than int16 indices)", i, (double)ext);
        g.size[i] = (int)ext;
        g.stride[i] = (i == 0) ? vd : g.stride[i - 1] * g.size[i - 1];
        cells *= g.size[i];
        if (cells > ((int64_t)1 << 34)) return fail(GT_ENOMEM, "bilateral grid of more than 2^34 cells");
    }
    gt::ValScale vsc{};
    for (int k = 0; k < vd; k++) {
        float vmax = 1.0f;
        if (k < nval) vmax = (vm[3 * k] <= vm[3 * k + 1]) ? std::max(std::fabs(vm[3 * k]), std::fabs(vm[3 * k + 1])) : 0.0f;
        const double bound = std::max(1e-30, (double)std::max(n, 1) * (double)vmax * 1.001);
        const int p = std::min(120, 62 - (int)std::ceil(std::log2(bound)));
        if (p < -120) return fail(GT_ERANGE, "values of channel %d are too large", k);
        vsc.s[k] = std::ldexp(1.0f, p);
        vsc.inv_s[k] = std::ldexp(1.0f, -p);
    }
    const size_t total = (size_t)cells * vd;
    GT_TRY(ensure(c->grid_acc, total * 8));
    GT_TRY(ensure(c->grid_a, total * 4));
    GT_TRY(ensure(c->grid_b, total * 4));
    if (nonfinite) {
        GT_TRY(ensure(c->grid_marks, total * 4));
        g.marks = (uint32_t *)c->grid_marks.p;
    }
    GT_HIP(gt::launch_grid(d, bilat, src, n, bilat ? src.W : 0, slice, n_slice, dst, g, cells, vsc,
                           (unsigned long long *)c->grid_acc.p, (float *)c->grid_a.p, (float *)c->grid_b.p,
                           c->profiling ? c->ev : nullptr, c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    if (st) {
        st->M = (int)std::min<int64_t>(cells, 0x7fffffff);
        st->capacity = st->M;
        if (c->profiling) {
            st->ms_splat = elapsed(c->ev[0], c->ev[1]);
            st->ms_sort = elapsed(c->ev[1], c->ev[2]);
            st->ms_nbr = 0.0f;
            st->ms_blur = elapsed(c->ev[3], c->ev[4]);
            st->ms_slice = elapsed(c->ev[4], c->ev[5]);
            st->ms_total = elapsed(c->ev[0], c->ev[5]);
        }
    }
    c->have = false;   // no lattice to export
    return GT_OK;
}

int jb_dense(gt_ctx *c, int W, int H, int F, int C, const float *im, int64_t im_cs, int Cr, const float *ref,
             int64_t ref_cs, float fw, float fh, float ff, float color_sigma, float *out, int64_t out_cs,
             int storage, const Tune &tu, gt_stats *st) {
    const bool fx = W > 1 && fw < 10 * W;
    const bool fy = H > 1 && fh < 10 * H;
    const bool ft = F > 1 && ff < 10 * F;
    const int d = Cr + (fx ? 1 : 0) + (fy ? 1 : 0) + (ft ? 1 : 0);
    const int vd = C + 1;
    GT_TRY(check_dims(d, vd));
    const int64_t n64 = (int64_t)W * H * F;
    if (n64 >= (1ll << 31)) return fail(GT_EINVAL, "image of %lld pixels exceeds 2^31", (long long)n64);
    const int n = (int)n64;
    gt::PixSrc src{};
    src.ref = ref;
    src.ref_cs = ref_cs;
    src.cr = Cr;
    src.W = W;
    src.H = H;
    src.F = F;
    src.fx = fx;
    src.fy = fy;
    src.ft = ft;
    src.ic = 1.0f / color_sigma;
    src.ix = fx ? 1.0f / fw : 0.f;
    src.iy = fy ? 1.0f / fh : 0.f;
    src.it = ft ? 1.0f / ff : 0.f;
    src.val = im;
    src.val_cs = im_cs;
    src.nval = C;
    src.homog = 1;
    SliceSpec sl;
    sl.enabled = true;
    sl.src = src;
    sl.n = n;
    sl.dst = gt::OutDst{out, out_cs, 1};
    if (tu.grid) return run_grid(c, src, n, d, vd, src, n, sl.dst, true, st);
    return run(c, src, n, W, d, vd, storage, tu, sl, false, st);
}

struct View {
    const float *p;
    int W, H, F;
    int64_t ys, ts, cs;
};

// JointBilateral::apply's zero-sigma recursion (src/GaussTransform.cpp:512-538)
// over strided views; dense views run the lattice directly, other views are
// gathered into a dense scratch image first.
int jb_view(gt_ctx *c, View im, int C, View ref, int Cr, View out, float fw, float fh, float ff, float cs,
            int storage, const Tune &tu, gt_stats *st) {
    if (im.W == 1) fw = 0;
    if (im.H == 1) fh = 0;
    if (im.F == 1) ff = 0;
    if (im.W > 1 && fw == 0) {
        for (int x = 0; x < im.W; x++) {
            View a = im, b = ref, o = out;
            a.p += x; b.p += x; o.p += x;
            a.W = b.W = o.W = 1;
            GT_TRY(jb_view(c, a, C, b, Cr, o, 0, fh, ff, cs, storage, tu, st));
        }
        return GT_OK;
    }
    if (im.H > 1 && fh == 0) {
        for (int y = 0; y < im.H; y++) {
            View a = im, b = ref, o = out;
            a.p += y * im.ys; b.p += y * ref.ys; o.p += y * out.ys;
            a.H = b.H = o.H = 1;
            GT_TRY(jb_view(c, a, C, b, Cr, o, fw, 0, ff, cs, storage, tu, st));
        }
        return GT_OK;
    }
    if (im.F > 1 && ff == 0) {
        for (int t = 0; t < im.F; t++) {
            View a = im, b = ref, o = out;
            a.p += t * im.ts; b.p += t * ref.ts; o.p += t * out.ts;
            a.F = b.F = o.F = 1;
            GT_TRY(jb_view(c, a, C, b, Cr, o, fw, fh, 0, cs, storage, tu, st));
        }
        return GT_OK;
    }
    const int W = im.W, H = im.H, F = im.F;
    const int64_t n = (int64_t)W * H * F;
    auto dense = [&](const View &v) { return v.ys == W && v.ts == (int64_t)W * H; };
    if (dense(im) && dense(ref) && dense(out))
        return jb_dense(c, W, H, F, C, im.p, im.cs, Cr, ref.p, ref.cs, fw, fh, ff, cs, (float *)out.p, out.cs,
                        storage, tu, st);
    GT_TRY(ensure(c->view_a, (size_t)n * C * 4));
    GT_TRY(ensure(c->view_b, (size_t)n * Cr * 4));
    GT_TRY(ensure(c->view_o, (size_t)n * C * 4));
    float *da = (float *)c->view_a.p, *db = (float *)c->view_b.p, *dob = (float *)c->view_o.p;
    GT_HIP(gt::launch_view_copy(im.p, im.ys, im.ts, im.cs, da, W, (int64_t)W * H, n, W, H, F, C, c->stream));
    GT_HIP(gt::launch_view_copy(ref.p, ref.ys, ref.ts, ref.cs, db, W, (int64_t)W * H, n, W, H, F, Cr, c->stream));
    GT_TRY(jb_dense(c, W, H, F, C, da, n, Cr, db, n, fw, fh, ff, cs, dob, n, storage, tu, st));
    GT_HIP(gt::launch_view_copy(dob, W, (int64_t)W * H, n, (float *)out.p, out.ys, out.ts, out.cs, W, H, F, C,
                                c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    return GT_OK;
}

} // namespace

namespace gt {
void invalidate_plan(gt_ctx *c) { c->plan_valid = false; }

gt_density_table *density_table_new() { return new gt_density_table(); }
void density_table_delete(gt_density_table *t) { delete t; }
void share_density(gt_ctx *c, gt_density_table *t) { c->shared_density = t; }
int ctx_device(gt_ctx *c) { return c->device; }
hipStream_t ctx_stream(gt_ctx *c) { return c->stream; }
} // namespace gt

extern "C" {

const char *gt_last_error(void) { return g_err; }

int gt_ctx_create(gt_ctx **out, int device) {
    if (!out) return fail(GT_EINVAL, "null context pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(GT_EDEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(GT_EINVAL, "device %d out of range (%d devices)", device, ndev);
    GT_HIP(hipSetDevice(device));
    gt_ctx *c = new gt_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(GT_EDEVICE, "stream creation failed");
    }
    c->stream = c->own;
    for (auto &e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            gt_ctx_destroy(c);
            return fail(GT_EDEVICE, "event creation failed");
        }
    }
    if (gt::alloc_should_fail() ||
        hipHostMalloc((void **)&c->h_ctr, sizeof(gt::Counters), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_mm, sizeof(float) * 6 * gt::kMaxDH, hipHostMallocDefault) != hipSuccess) {
        gt_ctx_destroy(c);
        return fail(GT_ENOMEM, "pinned allocation failed");
    }
    *out = c;
    return GT_OK;
}

void gt_ctx_destroy(gt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    if (c->h_batch) (void)hipHostFree(c->h_batch);
    free_buf(c->batch_desc);
    DevBuf *bufs[] = {&c->tkeys,       &c->tacc,      &c->tnf,      &c->vscale,    &c->vs_part,
                      &c->ctr,         &c->counts,    &c->slot_list, &c->slot_id,
                      &c->pk_a,        &c->pk_b,      &c->pay_b,    &c->hist,      &c->sort_status,
                      &c->sort_hist,   &c->vkeys,     &c->vals_a,   &c->vals_b,    &c->nbr,       &c->splat_copy,
                      &c->mm_part,     &c->mm_out,    &c->stage_in, &c->stage_out, &c->view_a,    &c->view_b,
                      &c->view_o,      &c->grid_acc,  &c->grid_a,   &c->grid_b,    &c->grid_marks,
                      &c->patch_w,     &c->patch_guide, &c->rp_rec, &c->rp_gs,
                      &c->rp_tent,     &c->rp_tcnt,   &c->tags};
    for (DevBuf *b : bufs) free_buf(*b);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->own) (void)hipStreamDestroy(c->own);
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    if (c->h_mm) (void)hipHostFree(c->h_mm);
    delete c;
}

int gt_ctx_set_stream(gt_ctx *c, void *stream) {
    if (!c) return fail(GT_EINVAL, "null context");
    c->stream = stream ? (hipStream_t)stream : c->own;
    return GT_OK;
}

int gt_ctx_set_profiling(gt_ctx *c, int level) {
    if (!c) return fail(GT_EINVAL, "null context");
    c->profiling = level < 0 ? 0 : level > 2 ? 2 : level;
    return GT_OK;
}

int gt_gauss_transform(gt_ctx *c, int n, int row_w, int d, int vd, const float *splat, const float *values,
                       int n_slice, const float *slice, const float *inv_sigma, float *out, gt_storage storage,
                       const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (vd > GT_MAX_VD) {
        // More value channels than one lattice carries (the reference takes
        // any number, src/Permutohedral.h:229-240): groups of GT_MAX_VD, each
        // a lattice over the same positions.  A channel's sums, blur and
        // slice never involve another channel, so every channel comes out as
        // it would from one lattice of all of them.
        if (n < 0 || n_slice < 0) return fail(GT_EINVAL, "negative pixel count");
        const int ns = (slice && slice != splat) ? n_slice : n;
        gt_stats total{};
        for (int c0 = 0; c0 < vd; c0 += GT_MAX_VD) {
            const int g = std::min(GT_MAX_VD, vd - c0);
            gt_stats gs{};
            GT_TRY(gt_gauss_transform(c, n, row_w, d, g, splat, values ? values + (size_t)c0 * n : nullptr, n_slice,
                                      slice, inv_sigma, out ? out + (size_t)c0 * ns : nullptr, storage, tune, &gs));
            add_stats(total, gs);
        }
        if (st) *st = total;
        return GT_OK;
    }
    GT_TRY(check_dims(d, vd));
    GT_TRY(check_storage(storage));
    if (n < 0 || n_slice < 0) return fail(GT_EINVAL, "negative pixel count");
    if (n > 0 && (!splat || !values || !inv_sigma)) return fail(GT_EINVAL, "null input pointer");
    if (slice == splat) slice = nullptr;
    const int ns = slice ? n_slice : n;
    if (ns > 0 && !out) return fail(GT_EINVAL, "null output pointer");
    if (row_w < 0 || (row_w > 0 && n % row_w != 0)) row_w = 0;
    GT_HIP(hipSetDevice(c->device));
    if (st) memset(st, 0, sizeof(*st));
    if (n == 0) {
        // No Gaussians: the transform is identically zero.
        if (ns > 0) GT_HIP(hipMemsetAsync(out, 0, (size_t)ns * vd * 4, c->stream));
        GT_HIP(hipStreamSynchronize(c->stream));
        return GT_OK;
    }
    const Tune tu = resolve(tune);
    gt::PixSrc src{};
    src.pos = splat;
    src.pos_cs = n;
    for (int i = 0; i < d; i++) src.inv[i] = inv_sigma[i];
    src.val = values;
    src.val_cs = n;
    src.nval = vd;
    SliceSpec sl;
    sl.enabled = ns > 0;
    sl.src = src;
    if (slice) {
        sl.src.pos = slice;
        sl.src.pos_cs = ns;
    }
    sl.n = ns;
    sl.dst = gt::OutDst{out, (int64_t)ns, 0};
    // The slice positions' tiling uses the splat row width only when they are the splat positions.
    return run(c, src, n, slice ? 0 : row_w, d, vd, storage, tu, sl, false, st);
}

int gt_gauss_transform_host(gt_ctx *c, int n, int row_w, int d, int vd, const float *splat, const float *values,
                            int n_slice, const float *slice, const float *inv_sigma, float *out,
                            gt_storage storage, const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(check_dims(d, std::min(vd, GT_MAX_VD)));   // (more channels run as groups)
    if (n < 0 || n_slice < 0) return fail(GT_EINVAL, "negative pixel count");
    const bool replay = (slice == nullptr || slice == splat);
    const int ns = replay ? n : n_slice;
    GT_HIP(hipSetDevice(c->device));
    const size_t in_floats = (size_t)n * (d + vd) + (replay ? 0 : (size_t)ns * d);
    GT_TRY(ensure(c->stage_in, in_floats * 4));
    GT_TRY(ensure(c->stage_out, (size_t)ns * vd * 4));
    float *dsp = (float *)c->stage_in.p;
    float *dva = dsp + (size_t)n * d;
    float *dsl = replay ? nullptr : dva + (size_t)n * vd;
    if (n) {
        GT_HIP(hipMemcpyAsync(dsp, splat, (size_t)n * d * 4, hipMemcpyHostToDevice, c->stream));
        GT_HIP(hipMemcpyAsync(dva, values, (size_t)n * vd * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (!replay && ns) GT_HIP(hipMemcpyAsync(dsl, slice, (size_t)ns * d * 4, hipMemcpyHostToDevice, c->stream));
    GT_TRY(gt_gauss_transform(c, n, row_w, d, vd, dsp, dva, ns, dsl, inv_sigma, (float *)c->stage_out.p, storage,
                              tune, st));
    if (ns) GT_HIP(hipMemcpy(out, c->stage_out.p, (size_t)ns * vd * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

static int jb_entry(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref, float fw,
                    float fh, float ff, float color_sigma, float *out, gt_storage storage, const gt_tuning *tune,
                    gt_stats *st, bool grid) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (W < 1 || H < 1 || F < 1 || C < 1 || Cr < 1) return fail(GT_EINVAL, "bad image size %dx%dx%d", W, H, F);
    if (!im || !ref || !out) return fail(GT_EINVAL, "null image pointer");
    GT_TRY(check_storage(storage));
    GT_HIP(hipSetDevice(c->device));
    if (st) memset(st, 0, sizeof(*st));
    Tune tu = resolve(tune);
    tu.grid = grid;
    const int64_t n = (int64_t)W * H * F;
    // Image channels beyond GT_MAX_VD - 1 (with the homogeneous one, the
    // values one lattice carries) run as groups against the same guide: a
    // channel's sums, blur, slice and normalisation never involve another
    // image channel, so each comes out as from one lattice of all of them.
    const int per = GT_MAX_VD - 1;
    gt_stats total{};
    for (int c0 = 0; c0 < C; c0 += per) {
        const int g = std::min(per, C - c0);
        View a{im + (size_t)c0 * n, W, H, F, W, (int64_t)W * H, n}, b{ref, W, H, F, W, (int64_t)W * H, n},
            o{out + (size_t)c0 * n, W, H, F, W, (int64_t)W * H, n};
        gt_stats gs{};
        GT_TRY(jb_view(c, a, g, b, Cr, o, fw, fh, ff, color_sigma, storage, tu, st ? &gs : nullptr));
        add_stats(total, gs);
    }
    if (st) *st = total;
    return GT_OK;
}

int gt_joint_bilateral(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref, float fw,
                       float fh, float ff, float color_sigma, float *out, gt_storage storage, const gt_tuning *tune,
                       gt_stats *st) {
    return jb_entry(c, W, H, F, C, im, Cr, ref, fw, fh, ff, color_sigma, out, storage, tune, st, false);
}

int gt_joint_bilateral_grid(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref,
                            float fw, float fh, float ff, float color_sigma, float *out, gt_stats *st) {
    return jb_entry(c, W, H, F, C, im, Cr, ref, fw, fh, ff, color_sigma, out, GT_STORE_F32, nullptr, st, true);
}

static int jb_host_entry(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref,
                         float fw, float fh, float ff, float color_sigma, float *out, gt_storage storage,
                         const gt_tuning *tune, gt_stats *st, bool grid) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (W < 1 || H < 1 || F < 1 || C < 1 || Cr < 1) return fail(GT_EINVAL, "bad image size %dx%dx%d", W, H, F);
    if (!im || !ref || !out) return fail(GT_EINVAL, "null image pointer");
    GT_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)W * H * F;
    GT_TRY(ensure(c->stage_in, n * (C + Cr) * 4));
    GT_TRY(ensure(c->stage_out, n * C * 4));
    float *di = (float *)c->stage_in.p, *dr = di + n * C;
    GT_HIP(hipMemcpyAsync(di, im, n * C * 4, hipMemcpyHostToDevice, c->stream));
    GT_HIP(hipMemcpyAsync(dr, ref, n * Cr * 4, hipMemcpyHostToDevice, c->stream));
    GT_TRY(jb_entry(c, W, H, F, C, di, Cr, dr, fw, fh, ff, color_sigma, (float *)c->stage_out.p, storage, tune, st,
                    grid));
    GT_HIP(hipMemcpy(out, c->stage_out.p, n * C * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

int gt_joint_bilateral_host(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref,
                            float fw, float fh, float ff, float color_sigma, float *out, gt_storage storage,
                            const gt_tuning *tune, gt_stats *st) {
    return jb_host_entry(c, W, H, F, C, im, Cr, ref, fw, fh, ff, color_sigma, out, storage, tune, st, false);
}

int gt_joint_bilateral_grid_host(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref,
                                 float fw, float fh, float ff, float color_sigma, float *out, gt_stats *st) {
    return jb_host_entry(c, W, H, F, C, im, Cr, ref, fw, fh, ff, color_sigma, out, GT_STORE_F32, nullptr, st, true);
}

int gt_gauss_transform_grid(gt_ctx *c, int n, int d, int vd, const float *splat, const float *values, int n_slice,
                            const float *slice, const float *inv_sigma, float *out, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(check_dims(d, vd));
    if (n < 0 || n_slice < 0) return fail(GT_EINVAL, "negative pixel count");
    if (n > 0 && (!splat || !values || !inv_sigma)) return fail(GT_EINVAL, "null input pointer");
    if (slice == splat) slice = nullptr;
    const int ns = slice ? n_slice : n;
    if (ns > 0 && !out) return fail(GT_EINVAL, "null output pointer");
    if (slice && n_slice > 0 && !inv_sigma) return fail(GT_EINVAL, "null inverse sigmas");
    GT_HIP(hipSetDevice(c->device));
    if (st) memset(st, 0, sizeof(*st));
    if (n == 0) {
        if (ns > 0) GT_HIP(hipMemsetAsync(out, 0, (size_t)ns * vd * 4, c->stream));
        GT_HIP(hipStreamSynchronize(c->stream));
        return GT_OK;
    }
    gt::PixSrc src{};
    src.pos = splat;
    src.pos_cs = n;
    for (int i = 0; i < d; i++) src.inv[i] = inv_sigma[i];
    src.val = values;
    src.val_cs = n;
    src.nval = vd;
    gt::PixSrc sl = src;
    if (slice) {
        sl.pos = slice;
        sl.pos_cs = ns;
    }
    return run_grid(c, src, n, d, vd, sl, ns, gt::OutDst{out, (int64_t)ns, 0}, false, st);
}

int gt_gauss_transform_grid_host(gt_ctx *c, int n, int d, int vd, const float *splat, const float *values,
                                 int n_slice, const float *slice, const float *inv_sigma, float *out,
                                 gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(check_dims(d, vd));
    if (n < 0 || n_slice < 0) return fail(GT_EINVAL, "negative pixel count");
    const bool replay = (slice == nullptr || slice == splat);
    const int ns = replay ? n : n_slice;
    GT_HIP(hipSetDevice(c->device));
    const size_t in_floats = (size_t)n * (d + vd) + (replay ? 0 : (size_t)ns * d);
    GT_TRY(ensure(c->stage_in, in_floats * 4));
    GT_TRY(ensure(c->stage_out, (size_t)ns * vd * 4));
    float *dsp = (float *)c->stage_in.p;
    float *dva = dsp + (size_t)n * d;
    float *dsl = replay ? nullptr : dva + (size_t)n * vd;
    if (n) {
        GT_HIP(hipMemcpyAsync(dsp, splat, (size_t)n * d * 4, hipMemcpyHostToDevice, c->stream));
        GT_HIP(hipMemcpyAsync(dva, values, (size_t)n * vd * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (!replay && ns) GT_HIP(hipMemcpyAsync(dsl, slice, (size_t)ns * d * 4, hipMemcpyHostToDevice, c->stream));
    GT_TRY(gt_gauss_transform_grid(c, n, d, vd, dsp, dva, ns, dsl, inv_sigma, (float *)c->stage_out.p, st));
    if (ns) GT_HIP(hipMemcpy(out, c->stage_out.p, (size_t)ns * vd * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

int gt_joint_bilateral_batch(gt_ctx *c, int count, const gt_bilateral_image *images, int C, int Cr,
                             gt_storage storage, const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (count < 0 || (count > 0 && !images)) return fail(GT_EINVAL, "bad image list");
    if (C < 1 || Cr < 1 || C + 1 > GT_MAX_VD) return fail(GT_EINVAL, "bad channel counts %d, %d", C, Cr);
    GT_TRY(check_storage(storage));
    GT_HIP(hipSetDevice(c->device));
    if (st) memset(st, 0, sizeof(*st));
    const Tune tu = resolve(tune);
    const int tw = (tu.tile_w == 16 || tu.tile_w == 64) ? tu.tile_w : 32;
    const int th = 256 / tw;
    const int d = Cr + 2, vd = C + 1;
    const bool tiled = d <= 6 && vd <= 4;
    std::vector<gt::BatchImage> imgs;
    std::vector<int> single;
    int64_t n = 0, ntiles = 0;
    for (int i = 0; i < count; i++) {
        const gt_bilateral_image &b = images[i];
        if (b.W < 1 || b.H < 1 || !b.im || !b.ref || !b.out) return fail(GT_EINVAL, "bad image %d", i);
        // The batch holds images whose lattice is (colour, x, y): both axes
        // filtered (src/GaussTransform.cpp:512-546); others go one by one.
        const bool plain = b.W > 1 && b.H > 1 && b.sigma_s > 0 && b.sigma_s < 10 * b.W && b.sigma_s < 10 * b.H;
        if (!tiled || !plain) {
            single.push_back(i);
            continue;
        }
        gt::BatchImage bi;
        zero(bi);
        bi.ref = b.ref;
        bi.im = b.im;
        bi.out = b.out;
        bi.W = b.W;
        bi.H = b.H;
        bi.ic = 1.0f / b.sigma_r;
        bi.ix = 1.0f / b.sigma_s;
        bi.iy = 1.0f / b.sigma_s;
        bi.tile0 = (int)ntiles;
        ntiles += (int64_t)((b.W + tw - 1) / tw) * ((b.H + th - 1) / th);
        n += (int64_t)b.W * b.H;
        imgs.push_back(bi);
    }
    if (n >= (1ll << 31) || ntiles >= (1ll << 31)) return fail(GT_EINVAL, "batch of %lld pixels exceeds 2^31", (long long)n);
    if (!imgs.empty()) {
        // The images' signature; when an earlier sorted call measured their
        // vertex counts, image i takes the vertex-share slice of [0, 2^24)
        // starting at its vertex offset, spread over its rows by the sum of
        // its keys' fields, -c_d = d sf y / sigma_s to within 2(d + 1).
        uint64_t sig = 1469598103934665603ull;
        auto mix = [&sig](const void *p, size_t len) {
            for (size_t q = 0; q < len; q++) sig = (sig ^ ((const unsigned char *)p)[q]) * 1099511628211ull;
        };
        mix(&Cr, sizeof(Cr));
        mix(&C, sizeof(C));
        for (const gt::BatchImage &bi : imgs) {
            mix(&bi.W, sizeof(bi.W));
            mix(&bi.H, sizeof(bi.H));
            mix(&bi.ic, sizeof(bi.ic));
            mix(&bi.ix, sizeof(bi.ix));
        }
        if (sig == 0) sig = 1;
        bool parts = tu.parts && c->batch_sig == sig && c->batch_verts.size() == imgs.size();
        uint64_t total = 0;
        if (parts)
            for (uint32_t v : c->batch_verts) total += v;
        parts = parts && total > 0;
        if (parts) {
            float sf[gt::kMaxDH];
            gt::plan_scale_factors(d, sf);
            const int margin = 2 * (d + 1);
            uint64_t before = 0;
            for (size_t q = 0; q < imgs.size(); q++) {
                gt::BatchImage &bi = imgs[q];
                const uint64_t v = c->batch_verts[q];
                const double range = std::ceil((double)d * sf[d - 1] * (double)(bi.H - 1) * bi.iy) + 2.0 * margin;
                const uint64_t width24 =
                    std::max<uint64_t>(1, (uint64_t)((double)(1 << 24) * (double)v / (double)total));
                bi.part_start24 =
                    (uint32_t)std::min<double>((double)(1 << 24) - 1, (double)(1 << 24) * (double)before / (double)total);
                bi.part_range = (int)std::min(range, (double)(1 << 24));
                bi.part_mulw = (uint64_t)((double)width24 * (double)(1 << 24) / (bi.part_range + 1.0));
                before += v;
            }
        }
        const int nb = (int)imgs.size();
        if (c->h_batch_cap < nb) {
            if (c->h_batch) (void)hipHostFree(c->h_batch);
            c->h_batch = nullptr;
            c->h_batch_cap = 0;
            if (gt::alloc_should_fail() ||
                hipHostMalloc((void **)&c->h_batch, sizeof(gt::BatchImage) * nb, hipHostMallocDefault) != hipSuccess)
                return fail(GT_ENOMEM, "pinned allocation failed");
            c->h_batch_cap = nb;
        }
        memcpy(c->h_batch, imgs.data(), sizeof(gt::BatchImage) * nb);
        GT_TRY(ensure(c->batch_desc, sizeof(gt::BatchImage) * nb));
        GT_HIP(hipMemcpyAsync(c->batch_desc.p, c->h_batch, sizeof(gt::BatchImage) * nb, hipMemcpyHostToDevice,
                              c->stream));
        BatchSpec b;
        b.ba.img = (const gt::BatchImage *)c->batch_desc.p;
        b.ba.count = nb;
        b.ntiles = (int)ntiles;
        b.Cr = Cr;
        b.C = C;
        zero(b.tmpl);
        b.tmpl.cr = Cr;
        b.tmpl.fx = 1;
        b.tmpl.fy = 1;
        b.tmpl.ft = 0;
        b.tmpl.F = 1;
        b.tmpl.nval = C;
        b.tmpl.homog = 1;
        b.tmpl.ref = imgs[0].ref;   // marks the bilateral form; per-image fields come from the descriptors
        zero(b.tmpl_dst);
        b.tmpl_dst.normalize = 1;
        b.sig = sig;
        b.parts = parts;
        Tune t2 = tu;
        t2.tile_w = tw;
        SliceSpec none;
        GT_TRY(run(c, b.tmpl, (int)n, 0, d, vd, storage, t2, none, false, st, &b, &imgs));
    }
    for (int i : single) {
        const gt_bilateral_image &b = images[i];
        GT_TRY(gt_joint_bilateral(c, b.W, b.H, 1, C, b.im, Cr, b.ref, b.sigma_s, b.sigma_s, 0.0f, b.sigma_r, b.out,
                                  storage, tune, nullptr));
    }
    return GT_OK;
}

int gt_build_lattice(gt_ctx *c, int n, int row_w, int d, int vd, const float *splat, const float *values,
                     const float *inv_sigma, gt_storage storage, const gt_tuning *tune, int *M_out) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(check_dims(d, vd));
    GT_TRY(check_storage(storage));
    if (n < 1 || !splat || !values || !inv_sigma) return fail(GT_EINVAL, "bad lattice input");
    if (row_w < 0 || (row_w > 0 && n % row_w != 0)) row_w = 0;
    GT_HIP(hipSetDevice(c->device));
    const Tune tu = resolve(tune);
    gt::PixSrc src{};
    src.pos = splat;
    src.pos_cs = n;
    for (int i = 0; i < d; i++) src.inv[i] = inv_sigma[i];
    src.val = values;
    src.val_cs = n;
    src.nval = vd;
    c->have = false;
    SliceSpec none;
    GT_TRY(run(c, src, n, row_w, d, vd, storage, tu, none, true, nullptr));
    if (M_out) *M_out = c->M;
    return GT_OK;
}

int gt_lattice_export(gt_ctx *c, int16_t *keys, int32_t *nbr, float *splat_vals, float *blur_vals) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (!c->have) return fail(GT_ESTATE, "no lattice built in this context");
    GT_HIP(hipSetDevice(c->device));
    const int M = c->M, d = c->d, vd = c->vd;
    if (M == 0) return GT_OK;
    const int m_max = (int)std::min<uint64_t>(c->cap, (uint64_t)c->last_n * (d + 1));
    GT_TRY(ensure(c->stage_out, std::max((size_t)M * d * 2, (size_t)M * vd * 4)));
    if (keys) {
        GT_HIP(gt::launch_unpack_keys(d, c->packed, c->plan, (const uint64_t *)c->vkeys.p, M,
                                      (int16_t *)c->stage_out.p, c->stream));
        GT_HIP(hipMemcpyAsync(keys, c->stage_out.p, (size_t)M * d * 2, hipMemcpyDeviceToHost, c->stream));
        GT_HIP(hipStreamSynchronize(c->stream));
    }
    if (nbr)
        for (int j = 0; j <= d; j++)
            GT_HIP(hipMemcpy(nbr + (size_t)j * M * 2, (const int2 *)c->nbr.p + (size_t)j * m_max, (size_t)M * 8,
                             hipMemcpyDeviceToHost));
    if (splat_vals) GT_HIP(hipMemcpy(splat_vals, c->splat_copy.p, (size_t)M * vd * 4, hipMemcpyDeviceToHost));
    if (blur_vals) {
        const unsigned *sbits = c->storage ? &((const gt::ValScale *)c->vscale.p)->vmax_bits[d][0] : nullptr;
        GT_HIP(gt::launch_store_to_f32(c->storage, c->vals_cur, sbits, M, vd, c->vs, (float *)c->stage_out.p,
                                       c->stream));
        GT_HIP(hipMemcpyAsync(blur_vals, c->stage_out.p, (size_t)M * vd * 4, hipMemcpyDeviceToHost, c->stream));
        GT_HIP(hipStreamSynchronize(c->stream));
    }
    return GT_OK;
}

// --- non-local means (NLMeans::apply / NLMeans3D::apply after the PCA) ------

static int patch_args(int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
                      const float *filters, const float *out) {
    if (W < 1 || H < 1 || F < 1 || C < 1) return fail(GT_EINVAL, "bad image size %dx%dx%dx%d", W, H, F, C);
    if ((int64_t)W * H * F >= (1ll << 31)) return fail(GT_EINVAL, "image of more than 2^31 pixels");
    if (F > 65535 || (H + 15) / 16 > 65535) return fail(GT_EINVAL, "more than 65535 frames or 2^20 rows");
    if (K < 1 || K > GT_MAX_PATCH_DIMS)
        return fail(GT_EINVAL, "patch dimensions K=%d outside [1, %d]", K, GT_MAX_PATCH_DIMS);
    if (fw < 1 || fh < 1 || ft < 1 || fw > 64 || fh > 64 || ft > 64)
        return fail(GT_EINVAL, "filter size %dx%dx%d outside [1, 64] per axis", fw, fh, ft);
    if (((int64_t)fw * fh * ft) % 2 != 1)   // src/Convolve.cpp:144-145
        return fail(GT_EINVAL, "filter must have odd size (%d %d %d)", fw, fh, ft);
    if (!im || !filters || !out) return fail(GT_EINVAL, "null pointer");
    return GT_OK;
}

// JointBilateral's AUTO choice (src/GaussTransform.cpp:512-564) for the call
// NLMeans makes, JointBilateral::apply(image, guide, s, s, INF, patch_sigma):
// 1 = GRID, 2 = PERMUTOHEDRAL, 0 = EXACT, GKDTREE or the zero-sigma split.
static int nlmeans_method(int W, int H, int F, int K, float s) {
    float sx = s, sy = s, st = INFINITY;
    if (W == 1) sx = 0;
    if (H == 1) sy = 0;
    if (F == 1) st = 0;
    if ((W > 1 && sx == 0) || (H > 1 && sy == 0)) return 0;
    const bool fx = W > 1 && sx < 10 * W, fy = H > 1 && sy < 10 * H, ft = F > 1 && st < 10 * F;
    const int pos = K + (fx ? 1 : 0) + (fy ? 1 : 0) + (ft ? 1 : 0);
    const int zx = fx ? ((int)(sx * 6 + 1) | 1) : 1, zy = fy ? ((int)(sy * 6 + 1) | 1) : 1;
    const int zt = ft ? ((int)(st * 6 + 1) | 1) : 1;
    if (zx * zy * zt < 16) return 0;
    if (pos <= 4) return 1;
    if (pos <= 10) return 2;
    return 0;
}

int gt_patch_project(gt_ctx *c, int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
                     const float *filters, float *out) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(patch_args(W, H, F, C, im, fw, fh, ft, K, filters, out));
    GT_HIP(hipSetDevice(c->device));
    GT_TRY(ensure(c->patch_w, (size_t)K * C * fw * fh * ft * 4));
    GT_HIP(gt::launch_patch_project(im, W, H, F, C, filters, fw, fh, ft, K, (float *)c->patch_w.p, out,
                                    c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    return GT_OK;
}

int gt_nlmeans(gt_ctx *c, int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
               const float *filters, float spatial_sigma, float patch_sigma, float *out, gt_storage storage,
               const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(patch_args(W, H, F, C, im, fw, fh, ft, K, filters, out));
    const int m = nlmeans_method(W, H, F, K, spatial_sigma);
    if (m == 0)
        return fail(GT_EUNSUPPORTED, "JointBilateral's AUTO method for this call is EXACT, GKDTREE or a "
                                     "per-row/column split (not on the GPU)");
    GT_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)W * H * F;
    GT_TRY(ensure(c->patch_w, (size_t)K * C * fw * fh * ft * 4));
    GT_TRY(ensure(c->patch_guide, n * K * 4));
    GT_HIP(gt::launch_patch_project(im, W, H, F, C, filters, fw, fh, ft, K, (float *)c->patch_w.p,
                                    (float *)c->patch_guide.p, c->stream));
    return jb_entry(c, W, H, F, C, im, K, (const float *)c->patch_guide.p, spatial_sigma, spatial_sigma, INFINITY,
                    patch_sigma, out, storage, tune, st, m == 1);
}

int gt_nlmeans_host(gt_ctx *c, int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
                    const float *filters, float spatial_sigma, float patch_sigma, float *out, gt_storage storage,
                    const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    GT_TRY(patch_args(W, H, F, C, im, fw, fh, ft, K, filters, out));
    GT_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)W * H * F, nf = (size_t)K * C * fw * fh * ft;
    GT_TRY(ensure(c->stage_in, (n * C + nf) * 4));
    GT_TRY(ensure(c->stage_out, n * C * 4));
    float *di = (float *)c->stage_in.p, *df = di + n * C;
    GT_HIP(hipMemcpyAsync(di, im, n * C * 4, hipMemcpyHostToDevice, c->stream));
    GT_HIP(hipMemcpyAsync(df, filters, nf * 4, hipMemcpyHostToDevice, c->stream));
    GT_TRY(gt_nlmeans(c, W, H, F, C, di, fw, fh, ft, K, df, spatial_sigma, patch_sigma, (float *)c->stage_out.p,
                      storage, tune, st));
    GT_HIP(hipMemcpy(out, c->stage_out.p, n * C * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

} // extern "C"

// --- no-allocation entry point ----------------------------------------------
//
// gt_joint_bilateral_noalloc runs gt_joint_bilateral's pipeline with every
// device buffer a fixed share of `scratch`, a file-scope device array, so the
// call allocates nothing (no hipMalloc/hipFree, no graph objects).  The
// caller's context lends its stream and pinned counters; its own buffers are
// set aside for the call and restored afterwards.  Vertex ids are taken in
// compaction order (no sort buffers) and the hash table has a fixed size.

namespace {

constexpr int kNaPixels = 1 << 20;       // pixels per call
constexpr int kNaMaxD = 5;               // lattice dimensions: one frame, guide of <= 3 channels, x, y
constexpr int kNaMaxC = 3;               // image channels (+ the homogeneous one)
constexpr uint32_t kNaCap = 1u << 21;    // hash-table slots
constexpr int kNaVerts = 1 << 20;        // lattice vertices (table load <= 0.5)
constexpr int kNaKeyWords = 2;           // wide keys of d <= 8
constexpr int kNaMaxDevices = 64;

constexpr size_t na_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaCtr = na_round(sizeof(gt::Counters));
constexpr size_t kNaKeys = na_round((size_t)kNaCap * kNaKeyWords * 8);
constexpr size_t kNaAcc = na_round((size_t)kNaCap * (kNaMaxC + 1) * 8);
constexpr size_t kNaMarks = na_round((size_t)kNaCap * 4);
constexpr size_t kNaSlotId = na_round((size_t)kNaCap * 4);
constexpr size_t kNaCounts = na_round((size_t)kNaCap / gt::kCompactSlots * 8);   // counts, then offsets
constexpr size_t kNaVkeys = na_round((size_t)kNaVerts * kNaKeyWords * 8);
constexpr size_t kNaVals = na_round((size_t)kNaVerts * 4 * 4);   // value stride 4, FP32 rows
constexpr size_t kNaNbr = na_round((size_t)(kNaMaxD + 1) * kNaVerts * 8);
constexpr size_t kNaScale = na_round(sizeof(gt::ValScale));
constexpr size_t kNaScalePart = na_round((size_t)1024 * (kNaMaxC + 1) * 3 * 4);
constexpr size_t kNaMmPart = na_round((size_t)1024 * gt::kMaxDH * 3 * 4);
constexpr size_t kNaMmOut = na_round((size_t)gt::kMaxDH * 3 * 4);
constexpr size_t kNaBytes = kNaCtr + kNaKeys + kNaAcc + kNaMarks + kNaSlotId + kNaCounts + kNaVkeys + 2 * kNaVals +
                            kNaNbr + kNaScale + kNaScalePart + kNaMmPart + kNaMmOut;

} // namespace

// The no-allocation entry point's device memory (about 208 MiB), one copy per
// GPU the library runs on, held until the process ends.  Host code uses the
// address hipGetSymbolAddress gives for the current device, never this name.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

struct NaDevice {
    std::mutex mu;          // calls on one device take turns (they share its scratch)
    void *ptr = nullptr;    // the device's copy of `scratch`
    size_t acc_zero = 0;    // what a complete call leaves zero (the next call then clears less)
    bool tnf_clean = false;
};
NaDevice g_na[kNaMaxDevices];

// The context's buffers that the pipeline of a one-frame joint bilateral with
// ids in compaction order touches, in scratch order.
DevBuf gt_ctx::*const kNaBufs[] = {&gt_ctx::ctr,    &gt_ctx::tkeys,  &gt_ctx::tacc,    &gt_ctx::tnf,
                                   &gt_ctx::slot_id, &gt_ctx::counts, &gt_ctx::vkeys,   &gt_ctx::vals_a,
                                   &gt_ctx::vals_b,  &gt_ctx::nbr,    &gt_ctx::vscale,  &gt_ctx::vs_part,
                                   &gt_ctx::mm_part, &gt_ctx::mm_out};
const size_t kNaShare[] = {kNaCtr,  kNaKeys,  kNaAcc,   kNaMarks,     kNaSlotId, kNaCounts, kNaVkeys,
                           kNaVals, kNaVals,  kNaNbr,   kNaScale,     kNaScalePart, kNaMmPart, kNaMmOut};
constexpr int kNaNumBufs = (int)(sizeof(kNaShare) / sizeof(kNaShare[0]));
static_assert(sizeof(kNaBufs) / sizeof(kNaBufs[0]) == kNaNumBufs, "one share per buffer");

// Swaps the context's buffers for shares of the device's scratch for the
// lifetime of the object, and back.
struct NaArena {
    gt_ctx *c;
    NaDevice &dev;
    DevBuf saved[kNaNumBufs];
    size_t saved_acc_zero;
    bool saved_tnf_clean;
    NaArena(gt_ctx *ctx, NaDevice &d) : c(ctx), dev(d) {
        saved_acc_zero = c->acc_zero;
        saved_tnf_clean = c->tnf_clean;
        unsigned char *p = (unsigned char *)dev.ptr;
        for (int i = 0; i < kNaNumBufs; i++) {
            saved[i] = c->*kNaBufs[i];
            (c->*kNaBufs[i]).p = p;
            (c->*kNaBufs[i]).cap = kNaShare[i];
            p += kNaShare[i];
        }
        c->acc_zero = dev.acc_zero;
        c->tnf_clean = dev.tnf_clean;
        g_no_alloc = true;
    }
    ~NaArena() {
        g_no_alloc = false;
        dev.acc_zero = c->acc_zero;
        dev.tnf_clean = c->tnf_clean;
        for (int i = 0; i < kNaNumBufs; i++) c->*kNaBufs[i] = saved[i];
        c->acc_zero = saved_acc_zero;
        c->tnf_clean = saved_tnf_clean;
        c->have = false;   // the lattice it built lived in the scratch
    }
};

} // namespace

int gt_joint_bilateral_noalloc(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref,
                               float fw, float fh, float ff, float color_sigma, float *out, gt_storage storage,
                               const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (W < 1 || H < 1 || F < 1 || C < 1 || Cr < 1) return fail(GT_EINVAL, "bad image size %dx%dx%d", W, H, F);
    if (!im || !ref || !out) return fail(GT_EINVAL, "null image pointer");
    GT_TRY(check_storage(storage));
    if ((int64_t)W * H * F > kNaPixels)
        return fail(GT_EINVAL, "the no-allocation entry point takes at most %d pixels (%lld given)", kNaPixels,
                    (long long)W * H * F);
    if (C > kNaMaxC) return fail(GT_EINVAL, "the no-allocation entry point takes at most %d channels", kNaMaxC);
    // JointBilateral's singleton rule (src/GaussTransform.cpp:512-514); a
    // zero spatial sigma left after it would split the image into pieces
    if (W == 1) fw = 0;
    if (H == 1) fh = 0;
    if (F == 1) ff = 0;
    if ((W > 1 && fw == 0) || (H > 1 && fh == 0) || (F > 1 && ff == 0))
        return fail(GT_EINVAL, "the no-allocation entry point filters whole images: no zero spatial sigma");
    const int d = Cr + (W > 1 && fw < 10 * W) + (H > 1 && fh < 10 * H) + (F > 1 && ff < 10 * F);
    if (d > kNaMaxD)
        return fail(GT_EINVAL, "the no-allocation entry point takes lattices of at most %d dimensions (%d)", kNaMaxD,
                    d);
    if (c->device < 0 || c->device >= kNaMaxDevices) return fail(GT_EINVAL, "device %d out of range", c->device);
    GT_HIP(hipSetDevice(c->device));
    NaDevice &dev = g_na[c->device];
    std::lock_guard<std::mutex> lk(dev.mu);
    if (!dev.ptr) {   // the first call on this device: the address of its copy
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(GT_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        }
        dev.ptr = p;
    }
    if (st) memset(st, 0, sizeof(*st));
    Tune tu = resolve(tune);
    tu.graph = false;         // instantiating a graph allocates
    tu.nbr = 2;               // ids in compaction order: no sort buffers
    tu.fixed_cap = kNaCap;
    tu.vert_limit = kNaVerts;
    NaArena arena(c, dev);
    const int64_t n = (int64_t)W * H * F;
    return jb_dense(c, W, H, F, C, im, n, Cr, ref, n, fw, fh, ff, color_sigma, out, n, storage, tu, st);
}
