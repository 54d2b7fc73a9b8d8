/*
 * gtlattice.h - C ABI of the MI355X (gfx950) permutohedral-lattice Gauss
 * transform library, libgtlattice.so.
 *
 * The library replaces the lattice stages of ImageStack's
 * GaussTransform::apply, PERMUTOHEDRAL case (src/GaussTransform.cpp:163-226,
 * driving src/Permutohedral.h) with hand-written HIP kernels:
 *   splat  : elevation, enclosing simplex, barycentric weights
 *            (src/Permutohedral.h:290-393), GPU open-addressing hash insert and
 *            exact 64-bit fixed-point accumulation of weighted values (integer
 *            atomics commute, so the sums do not depend on scheduling);
 *   sort   : compaction of the vertex keys and LSD radix sort, so vertex ids
 *            are the lexicographic rank of the key (deterministic);
 *   blur   : neighbour table, then d+1 Jacobi passes of [1/4 1/2 1/4]
 *            (src/Permutohedral.h:430-488);
 *   slice  : recompute simplex, gather, optional homogeneous normalisation
 *            (src/Permutohedral.h:405-427, src/GaussTransform.cpp:673-676).
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero GT_E* code on
 *     failure; gt_last_error() then returns a message (thread-local).  No
 *     function throws.  The C++ adapter turns a failure into ImageStack's
 *     panic(), matching the reference's error behaviour.
 *   - Images are dense planar float32: channel c of pixel p is at
 *     base[c * plane + p] with plane = W*H*F and p = (t*H + y)*W + x, the
 *     layout of a dense ImageStack Image (src/Image.h:34-37, :57-67).
 *   - Functions without the _host suffix take DEVICE pointers and run on the
 *     context's stream; they return after the results are complete.
 *   - Limits: 1 <= d <= GT_MAX_D, lattice key coordinates within int16 (the
 *     reference's key type, src/Permutohedral.h:28), N < 2^31.  Any number
 *     of value channels: more than GT_MAX_VD (a joint bilateral's image: more
 *     than GT_MAX_VD - 1) run as groups over the same positions, which gives
 *     every channel the output of one lattice of all of them (a channel's
 *     sums, blur and slice never involve another).
 *   - NaN and Inf values spread through the sums, blur and slice as in the
 *     reference's FP32 arithmetic; NaN/Inf positions are an error (GT_ERANGE).
 */
#ifndef GTLATTICE_H
#define GTLATTICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_MAX_D 16
#define GT_MAX_VD 16
#define GT_MAX_PATCH_DIMS 16

enum {
    GT_OK = 0,
    GT_EINVAL = 1,      /* bad argument (sizes, dims, null pointers)          */
    GT_ENOMEM = 2,      /* device allocation failed / over the memory budget  */
    GT_ERANGE = 3,      /* lattice key outside int16, or NaN/Inf position     */
    GT_EDEVICE = 4,     /* HIP runtime error                                  */
    GT_ESTATE = 5,      /* call out of order (e.g. export before build)       */
    GT_EUNSUPPORTED = 6, /* the reference would use a method not on the GPU    */
    GT_ECANCELED = 7,   /* a served request cancelled before it ran           */
    GT_ETIMEDOUT = 8    /* a served request not started by its deadline        */
};

/* Storage type of lattice values between splat, blur and slice.  Blur and
 * slice arithmetic is always FP32; splat sums are 64-bit fixed point. */
typedef enum { GT_STORE_F32 = 0, GT_STORE_BF16 = 1, GT_STORE_F16 = 2 } gt_storage;

/* Tuning knobs.  Zero fields select defaults (or the autotuner's choice). */
typedef struct {
    float load_factor;   /* target hash-table load, 0.05..0.7 (default 0.35)       */
    int waves_per_cu;    /* resident waves per CU of the tiled and neighbour kernels:
                            12..64 caps occupancy through LDS; 0 = no cap         */
    int dedup;           /* LDS tile dedup / pre-reduction: 0 auto, 1 on, -1 off     */
    int tile_w;          /* pixel-tile width of a workgroup: 16, 32, 64; 0 auto
                            (32, or 16 for lattices of 7 or more dimensions)    */
    int graph;           /* 0: replay the pipeline as a cached hipGraph; -1: don't  */
    int neighbours;      /* vertex ids and neighbour table: 0 auto (hash with row-band
                            table parts, else for d >= 11 or an expected lattice
                            of at most 2^20 vertices),
                            1 radix-sorted ids and a merge of the sorted keys,
                            2 compaction-order ids and hash lookups (no sort).  The
                            output is the same bit for bit; gt_build_lattice always
                            sorts (the exported lattice is in key order).          */
    int parts;           /* hash table in row bands, 8 or more, an eighth of them per
                            XCD's L2: 0 auto (one-frame bilateral forms with at
                            least 2^14 slots; Gauss transforms given row_w, of at
                            least 2^20 positions, split by their last position
                            channel), -1 off.  The output is the same.             */
} gt_tuning;

/* Per-call statistics (all optional). */
typedef struct {
    int M;               /* lattice vertices                                       */
    int capacity;        /* hash-table slots used                                  */
    int retries;         /* table regrowths in this call                           */
    int sort_passes;     /* radix passes run                                       */
    /* HIP-event times on the context's stream (profiling on):                   */
    float ms_splat;      /* the splat kernel alone                                 */
    float ms_sort;       /* compaction, radix sort, finalize                       */
    float ms_nbr;        /* the neighbour-table kernel                             */
    float ms_blur;       /* the d+1 blur-pass kernels                              */
    float ms_slice;      /* the slice kernel alone                                 */
    float ms_total;      /* first to last                                          */
    unsigned long long probes;    /* hash probes (slot visits), splat+slice+nbr    */
    unsigned long long inserts;   /* global insert attempts (after LDS dedup)      */
    unsigned long long atomics;   /* global atomic adds issued by splat            */
} gt_stats;

typedef struct gt_ctx gt_ctx;

/* Context: device buffers, stream, events.  One per host thread. */
int gt_ctx_create(gt_ctx **ctx, int device);
void gt_ctx_destroy(gt_ctx *ctx);
/* Use an existing hipStream_t (NULL = the context's own stream). */
int gt_ctx_set_stream(gt_ctx *ctx, void *hip_stream);
/* 0: off; 1: per-stage HIP-event times in gt_stats; 2: also count hash
 * probes, global inserts and atomics (slower).  Profiled calls run the
 * pipeline eagerly, launch by launch (a captured graph has no stage events). */
int gt_ctx_set_profiling(gt_ctx *ctx, int level);
const char *gt_last_error(void);

/*
 * GaussTransform::apply(slice, splat, values, sigmas, PERMUTOHEDRAL)
 * (src/GaussTransform.cpp:115-226).
 *   splat    : d planes of n pixels (device)     positions = splat * inv_sigma
 *   values   : vd planes of n pixels (device)
 *   slice    : d planes of n_slice pixels, or NULL to slice at the splat
 *              positions (the reference's replay slice, :196-207)
 *   inv_sigma: d host floats, 1.0f/sigma_c as the reference computes (:131)
 *   out      : vd planes of n_slice pixels (device), unnormalised, as the
 *              reference returns them (:194)
 *   row_w    : image width W (pixels per row) for 2-D tiling, or 0
 */
int gt_gauss_transform(gt_ctx *ctx, int n, int row_w, int d, int vd,
                       const float *splat, const float *values,
                       int n_slice, const float *slice, const float *inv_sigma,
                       float *out, gt_storage storage, const gt_tuning *tune,
                       gt_stats *stats);

/* Same with host pointers: copies in, runs, copies out. */
int gt_gauss_transform_host(gt_ctx *ctx, int n, int row_w, int d, int vd,
                            const float *splat, const float *values,
                            int n_slice, const float *slice, const float *inv_sigma,
                            float *out, gt_storage storage, const gt_tuning *tune,
                            gt_stats *stats);

/*
 * JointBilateral::apply(im, ref, fw, fh, ff, colorSigma, PERMUTOHEDRAL)
 * (src/GaussTransform.cpp:504-678) with the position image built on the
 * device: (ref_c * (1.0f/colorSigma), x * (1.0f/fw), y * (1.0f/fh),
 * t * (1.0f/ff)), values [im, 1], output im_c / homogeneous.  Zero spatial
 * sigmas filter columns/rows/frames independently (:516-538) as one batch.
 *   im  : C planes of W*H*F pixels (device); ref: Cr planes (device)
 *   out : C planes (device); may alias im.
 */
int gt_joint_bilateral(gt_ctx *ctx, int W, int H, int F, int C, const float *im,
                       int Cr, const float *ref, float fw, float fh, float ff,
                       float color_sigma, float *out, gt_storage storage,
                       const gt_tuning *tune, gt_stats *stats);

int gt_joint_bilateral_host(gt_ctx *ctx, int W, int H, int F, int C, const float *im,
                            int Cr, const float *ref, float fw, float fh, float ff,
                            float color_sigma, float *out, gt_storage storage,
                            const gt_tuning *tune, gt_stats *stats);

/*
 * gt_joint_bilateral without allocating: no hipMalloc, hipFree or graph
 * objects in the call.  No counterpart in the reference (which allocates per
 * call): it is for callers that cannot allocate -- serving loops that must
 * not touch the allocator, harnesses that forbid it.  Every device buffer is
 * a fixed share of one static device array of about 208 MiB (each GPU the
 * library runs on holds its own copy until the process ends); ctx lends its
 * stream and pinned counters, and its own buffers are left as they were.
 * Limits, each rejected with GT_EINVAL: W*H*F <= 2^20 pixels, C <= 3, a
 * lattice of at most 5 dimensions (a one-frame guide of <= 3 channels plus
 * x, y), no zero spatial sigma on an axis longer than 1.  A lattice of more
 * than 2^20 vertices returns GT_ENOMEM.  The output equals gt_joint_bilateral's
 * bit for bit.  Calls on one GPU share its array and take turns (a lock); the
 * array keeps the last call's contents, and each call clears what it needs.
 * tune: load_factor, neighbours and graph are fixed here; the rest applies.
 */
int gt_joint_bilateral_noalloc(gt_ctx *ctx, int W, int H, int F, int C, const float *im,
                               int Cr, const float *ref, float fw, float fh, float ff,
                               float color_sigma, float *out, gt_storage storage,
                               const gt_tuning *tune, gt_stats *stats);

/*
 * The bilateral grid (src/DenseGrid.h, 5 taps), the GRID method that
 * JointBilateral's AUTO choice picks for at most 4 position channels
 * (src/GaussTransform.cpp:553-564):
 *   gt_gauss_transform_grid: GaussTransform::apply(slice, splat, values,
 *     sigmas, GRID) (src/GaussTransform.cpp:227-297); arguments as
 *     gt_gauss_transform, at most 6 position channels.
 *   gt_joint_bilateral_grid: JointBilateral::apply(im, ref, fw, fh, ff,
 *     colorSigma, GRID) (src/GaussTransform.cpp:504-678); arguments as
 *     gt_joint_bilateral.
 * Grid cells, fractions and the blur follow the reference's FP32 (and its
 * double-precision row end) exactly; the splat sums are exact 64-bit fixed
 * point.  gt_stats: M = capacity = grid cells; ms_sort = the fixed-point
 * finish.
 */
int gt_gauss_transform_grid(gt_ctx *ctx, int n, int d, int vd, const float *splat,
                            const float *values, int n_slice, const float *slice,
                            const float *inv_sigma, float *out, gt_stats *stats);
int gt_gauss_transform_grid_host(gt_ctx *ctx, int n, int d, int vd, const float *splat,
                                 const float *values, int n_slice, const float *slice,
                                 const float *inv_sigma, float *out, gt_stats *stats);
int gt_joint_bilateral_grid(gt_ctx *ctx, int W, int H, int F, int C, const float *im,
                            int Cr, const float *ref, float fw, float fh, float ff,
                            float color_sigma, float *out, gt_stats *stats);
int gt_joint_bilateral_grid_host(gt_ctx *ctx, int W, int H, int F, int C, const float *im,
                                 int Cr, const float *ref, float fw, float fh, float ff,
                                 float color_sigma, float *out, gt_stats *stats);

/*
 * Non-local means: the lattice's second consumer in the reference.
 *
 * gt_patch_project replaces Convolve::apply(im, filters, Convolve::Zero,
 * Multiply::Inner) (src/Convolve.cpp:248-270, convolveSingle :139-173) for a
 * filter bank of K*C planes of fw x fh x ft (PatchPCA::apply's output,
 * src/Statistics.cpp:1710-1770, or PatchPCA3D's): output channel k is the sum
 * over image channels c of the zero-boundary convolution of channel c with
 * filter plane k*C + c.  Same summation order and roundings as the
 * reference, so the result is bit-identical to it.  1 <= K <= 16; the filter
 * volume must be odd.  im: C planes of W*H*F; out: K planes (device).
 *
 * gt_nlmeans replaces NLMeans::apply (src/GaussTransform.cpp:867-874) and
 * NLMeans3D::apply (:934-941) after their PatchPCA call: the guide is the
 * patch projection of im, then JointBilateral::apply(im, guide, spatial_sigma,
 * spatial_sigma, INF, patch_sigma) with JointBilateral's AUTO choice
 * (:553-564): the lattice for 5..10 position channels, the bilateral grid for
 * 4 or fewer; EXACT (small filters), GKDTREE (more than 10) and the zero-sigma
 * split return GT_EUNSUPPORTED.  The reference parses a method argument but
 * does not pass it on (:873), so there is none here.  out (C planes) may be im.
 * The PCA itself (random patch sampling with the host's rand()) stays with the
 * caller.
 */
int gt_patch_project(gt_ctx *ctx, int W, int H, int F, int C, const float *im, int fw, int fh,
                     int ft, int K, const float *filters, float *out);
int gt_nlmeans(gt_ctx *ctx, int W, int H, int F, int C, const float *im, int fw, int fh, int ft,
               int K, const float *filters, float spatial_sigma, float patch_sigma, float *out,
               gt_storage storage, const gt_tuning *tune, gt_stats *stats);
int gt_nlmeans_host(gt_ctx *ctx, int W, int H, int F, int C, const float *im, int fw, int fh,
                    int ft, int K, const float *filters, float spatial_sigma, float patch_sigma,
                    float *out, gt_storage storage, const gt_tuning *tune, gt_stats *stats);

/*
 * A batch of joint-bilateral filters: many images of uneven sizes through one
 * lattice (the image index is part of every vertex key, so the images never
 * mix), one graph launch for the whole pipeline.  Each image is filtered as
 * gt_joint_bilateral(ctx, W, H, 1, C, im, Cr, ref, sigma_s, sigma_s, 0,
 * sigma_r, out) would filter it; images that need the zero-sigma recursion
 * or a lattice beyond d = 6 / vd = 4 are filtered one by one.
 */
typedef struct {
    int W, H;
    const float *im;     /* C planes of W*H pixels (device)        */
    const float *ref;    /* Cr planes (device)                     */
    float *out;          /* C planes (device); may alias im        */
    float sigma_s;       /* spatial standard deviation (x and y)   */
    float sigma_r;       /* colour standard deviation              */
} gt_bilateral_image;

int gt_joint_bilateral_batch(gt_ctx *ctx, int count, const gt_bilateral_image *images,
                             int C, int Cr, gt_storage storage, const gt_tuning *tune,
                             gt_stats *stats);

/*
 * Autotuner: times the pipeline on these inputs over a grid of tunings
 * (tile_w 16/32/64 x load_factor 0.25/0.35/0.5 x waves_per_cu none/12/20,
 * plus the defaults; then the best of those with each neighbour form),
 * `reps` timed calls each after one untimed call, and
 * returns the tuning with the lowest median time.  Trial outputs go to a
 * scratch buffer; the inputs are only read.  No reference counterpart: the
 * reference has no tunable parameters.
 */
typedef struct {
    int trials;          /* tunings timed (including the defaults)               */
    float best_ms;       /* median ms per call of the returned tuning            */
    float default_ms;    /* median ms per call of the default tuning             */
} gt_tune_report;

int gt_autotune_joint_bilateral(gt_ctx *ctx, int W, int H, int F, int C, const float *im,
                                int Cr, const float *ref, float fw, float fh, float ff,
                                float color_sigma, gt_storage storage, int reps,
                                gt_tuning *best, gt_tune_report *report);
int gt_autotune_joint_bilateral_batch(gt_ctx *ctx, int count, const gt_bilateral_image *images,
                                      int C, int Cr, gt_storage storage, int reps,
                                      gt_tuning *best, gt_tune_report *report);

/*
 * Request loop (serving).  A server owns `slots` worker threads, each with
 * its own context (lattice buffers stay resident between requests) and its
 * own stream.  A request is a bilateral filter of a host image by itself
 * (Bilateral::apply, src/GaussTransform.cpp:730-734, PERMUTOHEDRAL) with
 * sigma_s in x and y.  Requests are admitted in arrival order when their HBM
 * footprint estimate fits the budget next to the running groups (or nothing
 * runs); a worker that frees up takes the oldest request plus every waiting
 * request of the same channel count and storage (up to max_batch requests
 * and batch_pixels pixels) as one batched lattice.  Workers run
 * concurrently, so groups overlap on the device.  Results are the same as
 * one gt_joint_bilateral call per request up to the batch's fixed-point
 * scale (relative differences below 1e-6).  No reference counterpart: the
 * reference filters one image per process.
 */
typedef struct {
    int slots;           /* worker threads / streams (default 4, at most 16)     */
    int max_batch;       /* requests per batched lattice (default 16)            */
    int batch_pixels;    /* pixels per batched lattice (default 4 Mi)            */
    size_t hbm_budget;   /* bytes of admitted footprints (default: half the free memory) */
    int warm_side;       /* >0: at start each worker filters a dense warm_side^2 RGB image, so its
                            buffers already fit requests up to that size (no growth while serving) */
} gt_server_config;

typedef struct {
    int W, H, C;                 /* C = 1 (grey) or 3 (RGB)                      */
    const float *im;             /* host: pixel (x,y) of channel c at
                                    im[c * plane_pitch + y * row_pitch + x]      */
    long long row_pitch;         /* floats between rows   (>= W)                 */
    long long plane_pitch;       /* floats between planes                        */
    float *out;                  /* host: dense C x H x W                        */
    float sigma_s, sigma_r;      /* spatial and colour standard deviations       */
    gt_storage storage;
    double deadline_ms;          /* > 0: not started within this many ms of the
                                    submit -> completes with GT_ETIMEDOUT, never run
                                    (a request that has started runs to the end)  */
} gt_request;

typedef struct {
    int status;                  /* GT_OK or the request's error code            */
    int batch;                   /* requests in the lattice it ran in            */
    double latency_ms;           /* submit -> result in host memory              */
    double queue_ms;             /* submit -> taken by a worker                  */
    size_t footprint;            /* admitted HBM footprint estimate (bytes)      */
} gt_request_result;

typedef struct gt_server gt_server;

int gt_server_create(gt_server **server, int device, const gt_server_config *config);
void gt_server_destroy(gt_server *server);
/* Queues a request (the pixels are read when a worker takes it: keep `im` and
 * `out` alive until gt_server_wait returns).  Never blocks. */
int gt_server_submit(gt_server *server, const gt_request *request, long long *ticket);
/* Blocks until the request is done; returns its status. */
int gt_server_wait(gt_server *server, long long ticket, gt_request_result *result);
/* Cancels a request that no worker has taken yet: it completes at once with
 * GT_ECANCELED (gt_server_wait still collects it) and its `out` is not
 * written.  GT_ESTATE if it is running or done (it then runs to the end),
 * GT_EINVAL for an unknown ticket. */
int gt_server_cancel(gt_server *server, long long ticket);
/* Open-loop replay: submits request i at arrival_ms[i] after the call starts
 * (one submitting thread), then waits for all of them. */
int gt_server_replay(gt_server *server, int count, const gt_request *requests,
                     const double *arrival_ms, gt_request_result *results);

/*
 * ImageStack .tmp files (src/FileTMP.cpp:75-233).  Header: five int32 (width,
 * height, frames, channels, type code); samples follow in channel, frame,
 * row, column order, the planar layout used throughout this library.  A type
 * code outside 0..9 is read as an old-style file (four-int header, float
 * samples), as the reference's loader does (:210-214).
 *
 * gt_tmp_info / gt_tmp_read / gt_tmp_write work on host memory and need no
 * GPU; they replace FileTMP::load / FileTMP::save.  gt_tmp_load / gt_tmp_save
 * move the image between the file and device memory, the samples crossing
 * PCIe in the file's own type and converted on the GPU.  Conversions are the
 * reference's C casts as its x86-64 build evaluates them: (float)sample on
 * load; on save, truncation toward zero through int32 (through int64 for
 * uint32), keeping the low bits.  dims receives width, height, frames,
 * channels; capacity is in floats.
 */
typedef enum {
    GT_TMP_F32 = 0, GT_TMP_F64 = 1, GT_TMP_U8 = 2, GT_TMP_I8 = 3, GT_TMP_U16 = 4,
    GT_TMP_I16 = 5, GT_TMP_U32 = 6, GT_TMP_I32 = 7, GT_TMP_U64 = 8, GT_TMP_I64 = 9
} gt_tmp_type;

int gt_tmp_info(const char *path, int *dims, int *type);
int gt_tmp_read(const char *path, float *out, size_t capacity, int *dims);
int gt_tmp_write(const char *path, const float *im, int W, int H, int F, int C, int type);
int gt_tmp_load(gt_ctx *ctx, const char *path, float *dev_out, size_t capacity, int *dims);
int gt_tmp_save(gt_ctx *ctx, const char *path, const float *dev_im, int W, int H, int F, int C,
                int type);

/*
 * A TCP front end that speaks ImageStack's own -send / -receive protocol
 * (TCPConnection::sendImage / recvImage, src/Network.cpp:161-196; Send::apply
 * and Receive::apply, src/NetworkOps.cpp:79-121): four uint32 (width,
 * height, frames, channels), then float32 samples in channel, frame, row,
 * column order; one image per connection.  Each received image is filtered
 * (GT_NET_BILATERAL: -bilateral sigma_r sigma_s sigma_s 0 permutohedral,
 * i.e. JointBilateral::apply(im, im, ...)) or returned as is (GT_NET_ECHO)
 * and sent back in the same format: on the same connection when reply_port
 * is 0, else on a new connection to the sender's address at reply_port,
 * which is what an unmodified client running
 *     ImageStack -load a.tmp -send HOST PORT -receive REPLY_PORT -save b.tmp
 * listens for.  Bilateral images of one frame and 1 or 3 channels go through
 * the serving loop (gt_server, configured by `serving`); other shapes run on
 * the front end's own context.  Latency is measured from the last byte
 * received to the last byte of the reply sent.  No reference counterpart for
 * the server itself: the reference's -receive takes one image per call.
 */
typedef enum { GT_NET_ECHO = 0, GT_NET_BILATERAL = 1 } gt_net_op;

typedef struct {
    int port;                    /* listen port; 0 = any free port             */
    int reply_port;              /* 0: reply on the same connection            */
    int op;                      /* gt_net_op                                  */
    float sigma_s, sigma_r;      /* bilateral spatial and colour sigmas        */
    gt_storage storage;
    long long max_pixels;        /* largest accepted W*H*F (default 2^26)      */
    gt_server_config serving;    /* the request loop behind GT_NET_BILATERAL   */
    long long max_bytes;         /* largest accepted image, W*H*F*C*4 bytes
                                    (default 1 GiB); larger: dropped, failed   */
    long long max_inflight;      /* bytes the handlers hold at once, input and
                                    output of every image being served (default
                                    4 GiB): a request that would pass it is
                                    refused (connection closed) before any
                                    allocation                                 */
    const char *bind_addr;       /* IPv4 address to listen on; NULL or "" =
                                    127.0.0.1 (this host only), "0.0.0.0" =
                                    every interface                            */
} gt_net_config;

typedef struct {
    long long received, replied, failed;
    double p50_ms, p99_ms;       /* over the last 100 000 replied images        */
    long long refused;           /* requests turned away by max_inflight        */
    long long inflight_bytes;    /* bytes the handlers hold now                 */
} gt_net_stats;

typedef struct gt_net_server gt_net_server;

int gt_net_server_create(gt_net_server **server, int device, const gt_net_config *config);
int gt_net_server_port(gt_net_server *server);
int gt_net_server_stats(gt_net_server *server, gt_net_stats *stats);
void gt_net_server_destroy(gt_net_server *server);

/*
 * Lattice inspection for parity checks: splat (and blur) the given points and
 * keep the lattice in the context.  Replaces reading
 * PermutohedralLattice::hashTable (src/Permutohedral.h:505-507).
 */
int gt_build_lattice(gt_ctx *ctx, int n, int row_w, int d, int vd,
                     const float *splat, const float *values, const float *inv_sigma,
                     gt_storage storage, const gt_tuning *tune, int *M_out);
/*
 * Copy the built lattice to host memory (any pointer may be NULL):
 *   keys      : M x d int16, sorted lexicographically (signed)
 *   nbr       : (d+1) x M x 2 int32, [axis][vertex] = {id(k+1-(d+1)e_j),
 *               id(k-1+(d+1)e_j)} or -1 (src/Permutohedral.h:446-451)
 *   splat_vals: M x vd float, values after splat (sorted order)
 *   blur_vals : M x vd float, values after the d+1 blur passes
 */
int gt_lattice_export(gt_ctx *ctx, int16_t *keys, int32_t *nbr,
                      float *splat_vals, float *blur_vals);

/*
 * Failure-injection hook for tests: of the library's next device and
 * pinned-host allocations, the first `after` succeed and the `count` after
 * them fail as if the memory were exhausted (count <= 0 turns it off).  A
 * call that meets one returns GT_ENOMEM and leaves its context (or server)
 * usable: the next call allocates again.  Process-wide.
 */
int gt_debug_fail_allocations(int after, int count);

#ifdef __cplusplus
}
#endif
#endif /* GTLATTICE_H */
