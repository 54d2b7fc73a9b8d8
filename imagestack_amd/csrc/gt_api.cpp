// C-ABI implementation of libgtlattice.so: context, device buffers, the
// splat -> sort -> blur -> slice pipeline, and the JointBilateral adapter.
// Host arithmetic that feeds the kernels (scale factors, reciprocals) is
// compiled with -ffp-contract=off and follows the reference's expressions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gt_internal.h"
#include "../../include/gtlattice.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define GT_HIP(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) return fail(GT_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define GT_TRY(x)                  \
    do {                           \
        int rc_ = (x);             \
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

int value_stride(int vd) { return vd <= 1 ? 1 : vd <= 2 ? 2 : vd <= 4 ? 4 : vd <= 8 ? 8 : 16; }
size_t store_bytes(int storage) { return storage == GT_STORE_F32 ? 4 : 2; }
int bitlen(uint32_t v) { int b = 0; while (v) { b++; v >>= 1; } return b; }

// Scale factors of src/Permutohedral.h:257-274, same expressions in FP32.
void scale_factors(int d, float *sf) {
    const float s23 = sqrtf((float)(2.0 / 3));
    for (int i = 0; i < d; i++) {
        float v = 1.0f / (sqrtf((float)(i + 1) * (float)(i + 2)));
        v *= (float)(d + 1) * s23;
        sf[i] = v;
    }
}

} // namespace

struct gt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr;
    int profiling = 0;   // 0 off, 1 HIP events, 2 events + probe/atomic counters
    hipEvent_t ev[8] = {};
    DevBuf tkeys, tvals, ctr, slot_list, slot_id, pk_a, pk_b, pay_b, hist, vkeys, vals_a, vals_b, nbr, splat_copy;
    DevBuf stage_in, stage_out, view_a, view_b, view_o;
    gt::Counters *h_ctr = nullptr;   // pinned
    // sizing memory: last lattice size seen for (n, d)
    int last_n = -1, last_d = -1;
    unsigned last_M = 0;
    // lattice kept for export
    bool have = false;
    int M = 0, d = 0, vd = 0, vs = 0, storage = 0;
    uint32_t cap = 0;
    void *vals_cur = nullptr;
};

namespace {

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return GT_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(GT_ENOMEM, "device allocation of %zu bytes failed", bytes);
    }
    b.cap = bytes;
    return GT_OK;
}

void free_buf(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

struct Tune {
    int block = 256;
    float lf = 0.35f;
    int dedup = -1;
    int tile_w = 32;
};

Tune resolve(const gt_tuning *t) {
    Tune r;
    if (!t) return r;
    if (t->block > 0) r.block = std::min(256, std::max(64, t->block & ~63));
    if (t->load_factor > 0) r.lf = std::min(0.7f, std::max(0.05f, t->load_factor));
    r.dedup = t->dedup;
    if (t->tile_w != 0) r.tile_w = t->tile_w;
    return r;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) { (void)hipGetLastError(); return 0.f; }
    return ms;
}

// Splat, sort, neighbour table and blur.  On return ctx holds the blurred
// lattice (ctx->vals_cur) and the table for the slice.
int build(gt_ctx *c, const gt::PixSrc &src, int n, int row_w, int d, int vd, int storage, const Tune &tu,
          gt::LatConst &lc, gt::TableArgs &ta, bool keep_splat, gt_stats *st) {
    hipStream_t s = c->stream;
    const int kw = (d + 3) / 4;
    const bool count = c->profiling >= 2;
    scale_factors(d, lc.sf);
    for (int i = d; i < gt::kMaxDH; i++) lc.sf[i] = 0.f;

    // Table capacity: from the last lattice of the same shape, else a
    // generous first guess; regrown (and the splat redone) if it fills.
    uint64_t est;
    if (c->last_n == n && c->last_d == d && c->last_M > 0) est = (uint64_t)(c->last_M * 1.05) + 64;
    else est = std::min<uint64_t>((uint64_t)n * (d + 1), std::max<uint64_t>(65536, (uint64_t)n));
    uint32_t cap = next_pow2((uint64_t)(est / tu.lf));
    const bool dedup = (tu.dedup < 0) ? (d <= 6 && vd <= 4) : (tu.dedup != 0);
    int retries = 0;
    while (true) {
        GT_TRY(ensure(c->tkeys, (size_t)cap * kw * 8));
        GT_TRY(ensure(c->tvals, (size_t)cap * vd * 4));
        GT_TRY(ensure(c->ctr, sizeof(gt::Counters)));
        GT_TRY(ensure(c->slot_list, (size_t)cap * 4));
        GT_TRY(ensure(c->slot_id, (size_t)cap * 4));
        ta.tkeys = (uint64_t *)c->tkeys.p;
        ta.tvals = (float *)c->tvals.p;
        ta.mask = cap - 1;
        ta.max_fill = (uint32_t)(cap * 0.75);
        ta.slot_id = (int *)c->slot_id.p;
        gt::Counters *ctr = (gt::Counters *)c->ctr.p;
        GT_HIP(gt::launch_clear(ta.tkeys, (size_t)cap * kw, ta.tvals, (size_t)cap * vd, ctr, s));
        if (c->profiling) GT_HIP(hipEventRecord(c->ev[0], s));
        GT_HIP(gt::launch_splat(d, src.ref != nullptr, dedup, src, n, row_w, tu.tile_w, tu.block, vd, lc, ta, ctr,
                                count, s));
        if (c->profiling) GT_HIP(hipEventRecord(c->ev[1], s));
        GT_HIP(gt::launch_compact(d, ta, cap, (uint32_t *)c->slot_list.p, ctr, s));
        GT_HIP(hipMemcpyAsync(c->h_ctr, ctr, sizeof(gt::Counters), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        if (c->h_ctr->flags & gt::kFlagBad)
            return fail(GT_ERANGE, "position is NaN/Inf or a lattice key leaves the int16 range");
        if (!(c->h_ctr->flags & gt::kFlagFull)) break;
        if (++retries > 8 || cap >= (1u << 31))
            return fail(GT_ENOMEM, "lattice does not fit a hash table of %u slots", cap);
        const double done = std::max(1u, c->h_ctr->done_pixels);
        const double proj = (double)c->h_ctr->created * ((double)n / done) * 1.25;
        cap = std::max<uint32_t>(cap * 2, next_pow2((uint64_t)(proj / tu.lf)));
    }
    const gt::Counters hc = *c->h_ctr;
    const int M = (int)hc.listed;
    c->last_n = n;
    c->last_d = d;
    c->last_M = (unsigned)M;

    // Radix sort of the packed keys: only the bits that vary.
    int umin[gt::kMaxDH], off[gt::kMaxDH], bits[gt::kMaxDH];
    int total = 0;
    for (int i = 0; i < d; i++) {
        umin[i] = M > 0 ? hc.kmin[i] : 0;
        bits[i] = M > 0 ? bitlen((uint32_t)(hc.kmax[i] - hc.kmin[i])) : 0;
        total += bits[i];
    }
    for (int i = d - 1, acc = 0; i >= 0; i--) { off[i] = acc; acc += bits[i]; }
    const int pw = std::max(1, (total + 63) / 64);
    GT_TRY(ensure(c->pk_a, (size_t)pw * M * 8));
    GT_TRY(ensure(c->pk_b, (size_t)pw * M * 8));
    GT_TRY(ensure(c->pay_b, (size_t)M * 4));
    const size_t he = gt::radix_hist_entries(M);
    GT_TRY(ensure(c->hist, he * 4));
    GT_HIP(gt::launch_pack(d, ta, (const uint32_t *)c->slot_list.p, M, umin, off, pw, (uint64_t *)c->pk_a.p, s));
    gt::SortBufs sb{(uint64_t *)c->pk_a.p, (uint64_t *)c->pk_b.p, (uint32_t *)c->slot_list.p,
                    (uint32_t *)c->pay_b.p, (uint32_t *)c->hist.p, he};
    bool in_b = false;
    int passes = 0;
    GT_HIP(gt::radix_sort(sb, M, pw, total, &in_b, &passes, s));
    const uint32_t *sorted = in_b ? sb.pay_b : sb.pay_a;

    const int vs = value_stride(vd);
    const size_t vbytes = (size_t)M * vs * store_bytes(storage);
    GT_TRY(ensure(c->vkeys, (size_t)M * kw * 8));
    GT_TRY(ensure(c->vals_a, vbytes));
    GT_TRY(ensure(c->vals_b, vbytes));
    float *splat_copy = nullptr;
    if (keep_splat) {
        GT_TRY(ensure(c->splat_copy, (size_t)M * vd * 4));
        splat_copy = (float *)c->splat_copy.p;
    }
    GT_HIP(gt::launch_finalize(d, storage, ta, sorted, M, vd, vs, (uint64_t *)c->vkeys.p, c->vals_a.p, splat_copy,
                               s));
    if (c->profiling) GT_HIP(hipEventRecord(c->ev[2], s));

    // Neighbour table, then the d+1 blur passes (ping-pong).
    GT_TRY(ensure(c->nbr, (size_t)(d + 1) * M * 8));
    int2 *nbr = (int2 *)c->nbr.p;
    GT_HIP(gt::launch_neighbors(d, ta, (const uint64_t *)c->vkeys.p, M, nbr, (gt::Counters *)c->ctr.p, count, s));
    if (c->profiling) GT_HIP(hipEventRecord(c->ev[3], s));
    void *cur = c->vals_a.p, *nxt = c->vals_b.p;
    for (int j = 0; j <= d; j++) {
        GT_HIP(gt::launch_blur(storage, vs, nbr + (size_t)j * M, M, cur, nxt, s));
        std::swap(cur, nxt);
    }
    if (c->profiling) GT_HIP(hipEventRecord(c->ev[4], s));

    c->have = true;
    c->M = M;
    c->d = d;
    c->vd = vd;
    c->vs = vs;
    c->storage = storage;
    c->cap = cap;
    c->vals_cur = cur;
    if (st) {
        st->M = M;
        st->capacity = (int)std::min<uint32_t>(cap, 0x7fffffff);
        st->retries = retries;
        st->sort_passes = passes;
    }
    return GT_OK;
}

int finish_stats(gt_ctx *c, gt_stats *st) {
    if (!st) return GT_OK;
    if (c->profiling) {
        GT_HIP(hipEventSynchronize(c->ev[5]));
        st->ms_splat = elapsed(c->ev[0], c->ev[1]);
        st->ms_sort = elapsed(c->ev[1], c->ev[2]);
        st->ms_nbr = elapsed(c->ev[2], c->ev[3]);
        st->ms_blur = elapsed(c->ev[3], c->ev[4]);
        st->ms_slice = elapsed(c->ev[4], c->ev[5]);
        st->ms_total = elapsed(c->ev[0], c->ev[5]);
        if (c->profiling < 2) return GT_OK;
        gt::Counters hc;
        GT_HIP(hipMemcpy(&hc, c->ctr.p, sizeof(hc), hipMemcpyDeviceToHost));
        st->probes = hc.probes;
        st->inserts = hc.inserts;
        st->atomics = hc.atomics;
    }
    return GT_OK;
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
    gt::LatConst lc;
    gt::TableArgs ta;
    GT_TRY(build(c, src, n, W, d, vd, storage, tu, lc, ta, false, st));
    gt::OutDst dst{out, out_cs, 1};
    GT_HIP(gt::launch_slice(d, true, storage, src, n, vd, c->vs, lc, ta, c->vals_cur, dst,
                            (gt::Counters *)c->ctr.p, c->profiling >= 2, c->stream));
    if (c->profiling) GT_HIP(hipEventRecord(c->ev[5], c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    return finish_stats(c, st);
}

} // namespace

namespace gt {
hipError_t launch_view_copy(const float *src, int64_t sys, int64_t sts, int64_t scs, float *dst, int64_t dys,
                            int64_t dts, int64_t dcs, int W, int H, int F, int C, hipStream_t st);
}

namespace {

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

extern "C" {

const char *gt_last_error(void) { return g_err.c_str(); }

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
            delete c;
            return fail(GT_EDEVICE, "event creation failed");
        }
    }
    if (hipHostMalloc((void **)&c->h_ctr, sizeof(gt::Counters), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return fail(GT_ENOMEM, "pinned allocation failed");
    }
    *out = c;
    return GT_OK;
}

void gt_ctx_destroy(gt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf *bufs[] = {&c->tkeys, &c->tvals, &c->ctr, &c->slot_list, &c->slot_id, &c->pk_a, &c->pk_b, &c->pay_b,
                      &c->hist, &c->vkeys, &c->vals_a, &c->vals_b, &c->nbr, &c->splat_copy, &c->stage_in,
                      &c->stage_out, &c->view_a, &c->view_b, &c->view_o};
    for (DevBuf *b : bufs) free_buf(*b);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->own) (void)hipStreamDestroy(c->own);
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    delete c;
}

int gt_ctx_set_stream(gt_ctx *c, void *stream) {
    if (!c) return fail(GT_EINVAL, "null context");
    c->stream = stream ? (hipStream_t)stream : c->own;
    return GT_OK;
}

int gt_ctx_set_profiling(gt_ctx *c, int enable) {
    if (!c) return fail(GT_EINVAL, "null context");
    c->profiling = enable < 0 ? 0 : enable > 2 ? 2 : enable;
    return GT_OK;
}

int gt_gauss_transform(gt_ctx *c, int n, int row_w, int d, int vd, const float *splat, const float *values,
                       int n_slice, const float *slice, const float *inv_sigma, float *out, gt_storage storage,
                       const gt_tuning *tune, gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
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
    src.homog = 0;
    gt::LatConst lc;
    gt::TableArgs ta;
    GT_TRY(build(c, src, n, row_w, d, vd, storage, tu, lc, ta, false, st));
    gt::PixSrc ssrc = src;
    if (slice) {
        ssrc.pos = slice;
        ssrc.pos_cs = ns;
    }
    gt::OutDst dst{out, (int64_t)ns, 0};
    GT_HIP(gt::launch_slice(d, false, storage, ssrc, ns, vd, c->vs, lc, ta, c->vals_cur, dst,
                            (gt::Counters *)c->ctr.p, c->profiling >= 2, c->stream));
    if (c->profiling) GT_HIP(hipEventRecord(c->ev[5], c->stream));
    GT_HIP(hipStreamSynchronize(c->stream));
    GT_TRY(finish_stats(c, st));
    // A slice at new positions may fail the range checks too.
    gt::Counters hc;
    GT_HIP(hipMemcpy(&hc, c->ctr.p, sizeof(hc), hipMemcpyDeviceToHost));
    if (hc.flags & gt::kFlagBad)
        return fail(GT_ERANGE, "slice position is NaN/Inf or a lattice key leaves the int16 range");
    return GT_OK;
}

int gt_gauss_transform_host(gt_ctx *c, int n, int row_w, int d, int vd, const float *splat, const float *values,
                            int n_slice, const float *slice, const float *inv_sigma, float *out,
                            gt_storage storage, const gt_tuning *tune, gt_stats *st) {
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
    GT_TRY(gt_gauss_transform(c, n, row_w, d, vd, dsp, dva, ns, dsl, inv_sigma, (float *)c->stage_out.p, storage,
                              tune, st));
    if (ns) GT_HIP(hipMemcpy(out, c->stage_out.p, (size_t)ns * vd * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

int gt_joint_bilateral(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref, float fw,
                       float fh, float ff, float color_sigma, float *out, gt_storage storage, const gt_tuning *tune,
                       gt_stats *st) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (W < 1 || H < 1 || F < 1 || C < 1 || Cr < 1) return fail(GT_EINVAL, "bad image size %dx%dx%d", W, H, F);
    if (!im || !ref || !out) return fail(GT_EINVAL, "null image pointer");
    if (C + 1 > GT_MAX_VD) return fail(GT_EINVAL, "too many channels (%d)", C);
    GT_TRY(check_storage(storage));
    GT_HIP(hipSetDevice(c->device));
    if (st) memset(st, 0, sizeof(*st));
    const Tune tu = resolve(tune);
    const int64_t n = (int64_t)W * H * F;
    View a{im, W, H, F, W, (int64_t)W * H, n}, b{ref, W, H, F, W, (int64_t)W * H, n}, o{out, W, H, F, W,
                                                                                       (int64_t)W * H, n};
    return jb_view(c, a, C, b, Cr, o, fw, fh, ff, color_sigma, storage, tu, st);
}

int gt_joint_bilateral_host(gt_ctx *c, int W, int H, int F, int C, const float *im, int Cr, const float *ref,
                            float fw, float fh, float ff, float color_sigma, float *out, gt_storage storage,
                            const gt_tuning *tune, gt_stats *st) {
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
    GT_TRY(gt_joint_bilateral(c, W, H, F, C, di, Cr, dr, fw, fh, ff, color_sigma, (float *)c->stage_out.p, storage,
                              tune, st));
    GT_HIP(hipMemcpy(out, c->stage_out.p, n * C * 4, hipMemcpyDeviceToHost));
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
    gt::LatConst lc;
    gt::TableArgs ta;
    c->have = false;
    GT_TRY(build(c, src, n, row_w, d, vd, storage, tu, lc, ta, true, nullptr));
    GT_HIP(hipStreamSynchronize(c->stream));
    if (M_out) *M_out = c->M;
    return GT_OK;
}

int gt_lattice_export(gt_ctx *c, int16_t *keys, int32_t *nbr, float *splat_vals, float *blur_vals) {
    if (!c) return fail(GT_EINVAL, "null context");
    if (!c->have) return fail(GT_ESTATE, "no lattice built in this context");
    GT_HIP(hipSetDevice(c->device));
    const int M = c->M, d = c->d, vd = c->vd;
    if (M == 0) return GT_OK;
    GT_TRY(ensure(c->stage_out, std::max((size_t)M * d * 2, (size_t)M * vd * 4)));
    if (keys) {
        GT_HIP(gt::launch_unpack_keys(d, (const uint64_t *)c->vkeys.p, M, (int16_t *)c->stage_out.p, c->stream));
        GT_HIP(hipMemcpyAsync(keys, c->stage_out.p, (size_t)M * d * 2, hipMemcpyDeviceToHost, c->stream));
        GT_HIP(hipStreamSynchronize(c->stream));
    }
    if (nbr) GT_HIP(hipMemcpy(nbr, c->nbr.p, (size_t)(d + 1) * M * 8, hipMemcpyDeviceToHost));
    if (splat_vals) GT_HIP(hipMemcpy(splat_vals, c->splat_copy.p, (size_t)M * vd * 4, hipMemcpyDeviceToHost));
    if (blur_vals) {
        GT_HIP(gt::launch_store_to_f32(c->storage, c->vals_cur, M, vd, c->vs, (float *)c->stage_out.p, c->stream));
        GT_HIP(hipMemcpyAsync(blur_vals, c->stage_out.p, (size_t)M * vd * 4, hipMemcpyDeviceToHost, c->stream));
        GT_HIP(hipStreamSynchronize(c->stream));
    }
    return GT_OK;
}

} // extern "C"
