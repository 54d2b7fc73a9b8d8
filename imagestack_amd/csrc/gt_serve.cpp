// Request loop for bilateral filtering: requests (host images) are admitted
// by an HBM footprint estimate, grouped into batched lattices (continuous
// batching: whatever compatible requests are waiting when a worker frees up
// joins its batch), and run by `slots` worker threads, each with its own
// context (resident lattice buffers) and its own stream, so the groups
// overlap on the device.  Latency runs from submit to the result in host
// memory.  A request still queued can be cancelled, and one given a
// deadline that no worker takes in time completes with GT_ETIMEDOUT; both
// leave its output untouched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gt_internal.h"
#include "../../include/gtlattice.h"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

uint64_t pow2_at_least(uint64_t v) {
    uint64_t c = 1024;
    while (c < v) c <<= 1;
    return c;
}

// Device bytes one request occupies while it runs: its input and output
// planes plus the lattice buffers sized as the library sizes them for an
// unseen image (hash table for min(n(d+1), max(65536, n)) vertices at load
// 0.35; vertex arrays for that many vertices; the splat's replay record over
// the image padded to whole tiles of up to 64 x 16 pixels).
size_t footprint(const gt_request &r) {
    const uint64_t n = (uint64_t)r.W * r.H;
    const int d = r.C + 2, vd = r.C + 1, vs = vd <= 2 ? 2 : 4;
    const uint64_t est = std::min<uint64_t>(n * (d + 1), std::max<uint64_t>(65536, n));
    const uint64_t cap = pow2_at_least((uint64_t)(est / 0.35));
    const uint64_t m = std::min<uint64_t>(cap, n * (d + 1));
    const size_t sb = r.storage == GT_STORE_F32 ? 4 : 2;
    const uint64_t np = (uint64_t)((r.W + 63) / 64 * 64) * (uint64_t)((r.H + 15) / 16 * 16);
    return (size_t)(2 * r.C * n * 4                       // planes in / out
                    + cap * (8 + 8 * vd + 4)              // keys, fixed-point sums, slot ids
                    + m * (4 + 16 + 4 + 8 + 2 * vs * sb + 8 * (d + 1))    // compaction, sort, values, nbr
                    + np * (16 + 4 * (d + 1)) + np / 256 * (8 * 512 + 4));   // replay entries, tile lists
}

struct Req {
    gt_request r{};
    long long ticket = 0;
    size_t fp = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = GT_OK, batch = 0;
    bool taken = false;   // a worker runs it (no longer cancellable)
    bool done = false;
    std::string err;
    bool expired(double t) const { return r.deadline_ms > 0 && t - t_submit >= r.deadline_ms; }
};

} // namespace

struct gt_server {
    int device = 0;
    gt_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> queue;
    std::unordered_map<long long, Req *> reqs;
    long long next_ticket = 1;
    size_t reserved = 0;
    int running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::string init_err;
    gt_density_table *density = nullptr;   // vertex densities the workers share
    int init_rc = GT_OK;
    std::vector<float> warm;   // host warm-up image (3 planes of warm_side^2), shared read-only
};

namespace {

// Takes the next group under the lock: the oldest request (admitted when its
// footprint fits the budget next to what runs, or when nothing runs), then
// every other waiting request with the same channel count and storage while
// the group stays within max_batch requests, batch_pixels pixels and the
// budget.
std::vector<Req *> take_group(gt_server *s) {
    std::vector<Req *> g;
    Req *head = s->queue.front();
    head->taken = true;
    s->queue.pop_front();
    g.push_back(head);
    size_t fp = head->fp;
    uint64_t px = (uint64_t)head->r.W * head->r.H;
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)g.size() < s->cfg.max_batch;) {
        Req *q = *it;
        const uint64_t qpx = (uint64_t)q->r.W * q->r.H;
        if (q->r.C == head->r.C && q->r.storage == head->r.storage && px + qpx <= (uint64_t)s->cfg.batch_pixels &&
            s->reserved + fp + q->fp <= s->cfg.hbm_budget) {
            q->taken = true;
            g.push_back(q);
            fp += q->fp;
            px += qpx;
            it = s->queue.erase(it);
        } else {
            ++it;
        }
    }
    s->reserved += fp;
    s->running++;
    return g;
}

// Completes a queued request without running it (under the lock).
void finish_unrun(Req *q, int status, const char *msg, double t) {
    q->status = status;
    q->err = msg;
    q->t_start = q->t_done = t;
    q->done = true;
}

// Completes every queued request past its deadline with GT_ETIMEDOUT (under
// the lock); true if any was.
bool expire_queued(gt_server *s, double t) {
    bool any = false;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if ((*it)->expired(t)) {
            finish_unrun(*it, GT_ETIMEDOUT, "server: request not started before its deadline", t);
            it = s->queue.erase(it);
            any = true;
        } else {
            ++it;
        }
    }
    return any;
}

bool can_take(gt_server *s) {
    if (s->queue.empty()) return false;
    return s->running == 0 || s->reserved + s->queue.front()->fp <= s->cfg.hbm_budget;
}

struct DevArena {
    float *p = nullptr;
    size_t cap = 0;
    int ensure(size_t floats) {
        if (floats <= cap) return GT_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (gt::alloc_should_fail()) return GT_ENOMEM;
        if (hipMalloc((void **)&p, floats * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            return GT_ENOMEM;
        }
        cap = floats;
        return GT_OK;
    }
    ~DevArena() {
        if (p) (void)hipFree(p);
    }
};

// Pinned host staging of one worker, for requests whose pixels live in
// pageable memory: a 2-D copy from pageable memory goes row by row through
// the runtime's own bounce buffer (a 512^2 RGB upload took ~14 ms that way),
// so such planes are packed here by the CPU and cross PCIe in one copy.
struct HostStage {
    float *p = nullptr;
    size_t cap = 0;
    int ensure(size_t floats) {
        if (floats <= cap) return GT_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (gt::alloc_should_fail()) return GT_ENOMEM;
        if (hipHostMalloc((void **)&p, floats * sizeof(float), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return GT_ENOMEM;
        }
        cap = floats;
        return GT_OK;
    }
    ~HostStage() {
        if (p) (void)hipHostFree(p);
    }
};

bool is_pinned(const void *ptr) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

// Runs one group on the worker's context and stream; returns a status per
// request in `rc` and a message in `err`.
void run_group(gt_ctx *ctx, hipStream_t st, DevArena &arena, HostStage &stage, const std::vector<Req *> &g,
               std::vector<int> &rc, std::vector<std::string> &err) {
    const int C = g[0]->r.C;
    const gt_storage storage = g[0]->r.storage;
    size_t total = 0;
    for (Req *q : g) total += (size_t)q->r.W * q->r.H * C;
    auto fail_all = [&](int code, const std::string &m) {
        for (size_t i = 0; i < g.size(); i++) {
            rc[i] = code;
            err[i] = m;
        }
    };
    if (arena.ensure(2 * total) != GT_OK) return fail_all(GT_ENOMEM, "server: no device memory for the planes");
    float *din = arena.p, *dout = arena.p + total;
    std::vector<char> in_pinned(g.size()), out_pinned(g.size());
    bool staged = false;
    for (size_t i = 0; i < g.size(); i++) {
        in_pinned[i] = is_pinned(g[i]->r.im);
        out_pinned[i] = is_pinned(g[i]->r.out);
        staged = staged || !in_pinned[i] || !out_pinned[i];
    }
    if (staged && stage.ensure(total) != GT_OK) return fail_all(GT_ENOMEM, "server: no pinned host staging");
    size_t off = 0;
    std::vector<gt_bilateral_image> imgs;
    for (size_t i = 0; i < g.size(); i++) {
        const gt_request &r = g[i]->r;
        const size_t plane = (size_t)r.W * r.H;
        hipError_t e = hipSuccess;
        if (in_pinned[i]) {
            for (int c = 0; c < C && e == hipSuccess; c++)
                e = hipMemcpy2DAsync(din + off + c * plane, (size_t)r.W * 4, r.im + (size_t)c * r.plane_pitch,
                                     (size_t)r.row_pitch * 4, (size_t)r.W * 4, (size_t)r.H, hipMemcpyHostToDevice, st);
        } else {
            float *h = stage.p + off;
            for (int c = 0; c < C; c++)
                for (int y = 0; y < r.H; y++)
                    memcpy(h + c * plane + (size_t)y * r.W, r.im + (size_t)c * r.plane_pitch + (size_t)y * r.row_pitch,
                           (size_t)r.W * 4);
            e = hipMemcpyAsync(din + off, h, plane * C * 4, hipMemcpyHostToDevice, st);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail_all(GT_EDEVICE, std::string("server: upload failed: ") + hipGetErrorString(e));
        }
        imgs.push_back(gt_bilateral_image{r.W, r.H, din + off, din + off, dout + off, r.sigma_s, r.sigma_r});
        off += plane * C;
    }
    gt::invalidate_plan(ctx);
    gt_tuning tu{};
    tu.graph = -1;   // every group differs in shape: no graph to reuse
    int status;
    if (g.size() == 1) {
        const gt_request &r = g[0]->r;
        status = gt_joint_bilateral(ctx, r.W, r.H, 1, C, din, C, din, r.sigma_s, r.sigma_s, 0.0f, r.sigma_r, dout,
                                    storage, &tu, nullptr);
    } else {
        status = gt_joint_bilateral_batch(ctx, (int)imgs.size(), imgs.data(), C, C, storage, &tu, nullptr);
    }
    if (status != GT_OK) return fail_all(status, gt_last_error());
    off = 0;
    for (size_t i = 0; i < g.size(); i++) {
        const gt_request &r = g[i]->r;
        const size_t bytes = (size_t)r.W * r.H * C * 4;
        // the input staging of a request is free again once its upload ran: reuse it for the download
        void *dst = out_pinned[i] ? (void *)r.out : (void *)(stage.p + off);
        const hipError_t e = hipMemcpyAsync(dst, dout + off, bytes, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail_all(GT_EDEVICE, std::string("server: download failed: ") + hipGetErrorString(e));
        }
        off += (size_t)r.W * r.H * C;
    }
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail_all(GT_EDEVICE, std::string("server: stream failed: ") + hipGetErrorString(e));
    }
    off = 0;
    for (size_t i = 0; i < g.size(); i++) {
        const size_t n = (size_t)g[i]->r.W * g[i]->r.H * C;
        if (!out_pinned[i]) memcpy(g[i]->r.out, stage.p + off, n * 4);
        off += n;
    }
    for (size_t i = 0; i < g.size(); i++) rc[i] = GT_OK;
}

void worker_main(gt_server *s, std::mutex *init_mu, int *init_left, std::condition_variable *init_cv) {
    gt_ctx *ctx = nullptr;
    hipStream_t st = nullptr;
    int rc = hipSetDevice(s->device) == hipSuccess ? GT_OK : GT_EDEVICE;
    if (rc == GT_OK) rc = gt_ctx_create(&ctx, s->device);
    if (rc == GT_OK) gt::share_density(ctx, s->density);   // a filter learnt by one worker serves all
    if (rc == GT_OK && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = GT_EDEVICE;
    if (rc == GT_OK) rc = gt_ctx_set_stream(ctx, st);
    DevArena arena;
    HostStage stage;
    // Warm-up: one dense RGB filter of the configured size grows the context's
    // lattice buffers and the arena to their serving size.
    if (rc == GT_OK && s->cfg.warm_side > 0) {
        const int side = s->cfg.warm_side;
        const size_t n = (size_t)side * side;
        if (arena.ensure(6 * n) != GT_OK || stage.ensure(3 * n) != GT_OK) {
            rc = GT_ENOMEM;
        } else if (hipMemcpyAsync(arena.p, s->warm.data(), 3 * n * sizeof(float), hipMemcpyHostToDevice, st) !=
                   hipSuccess) {
            (void)hipGetLastError();
            rc = GT_EDEVICE;
        } else {
            gt_tuning tu{};
            tu.graph = -1;
            rc = gt_joint_bilateral(ctx, side, side, 1, 3, arena.p, 3, arena.p, 4.0f, 4.0f, 0.0f, 0.1f,
                                    arena.p + 3 * n, GT_STORE_F32, &tu, nullptr);
            gt::invalidate_plan(ctx);
        }
    }
    {
        std::lock_guard<std::mutex> lk(*init_mu);
        if (rc != GT_OK && s->init_rc == GT_OK) {
            s->init_rc = rc;
            s->init_err = gt_last_error();
        }
        (*init_left)--;
    }
    init_cv->notify_all();
    for (;;) {
        std::vector<Req *> g;
        bool expired = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            while (!s->stop) {
                s->cv_work.wait(lk, [&] { return s->stop || (rc == GT_OK && can_take(s)); });
                if (s->stop) break;
                expired = expire_queued(s, now_ms()) || expired;
                if (can_take(s)) {
                    g = take_group(s);
                    break;
                }
            }
        }
        if (expired) s->cv_done.notify_all();
        if (g.empty()) break;   // stopping
        const double t0 = now_ms();
        std::vector<int> st_rc(g.size(), GT_OK);
        std::vector<std::string> st_err(g.size());
        run_group(ctx, st, arena, stage, g, st_rc, st_err);
        const double t1 = now_ms();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            size_t fp = 0;
            for (size_t i = 0; i < g.size(); i++) {
                Req *q = g[i];
                q->t_start = t0;
                q->t_done = t1;
                q->status = st_rc[i];
                q->err = st_err[i];
                q->batch = (int)g.size();
                q->done = true;
                fp += q->fp;
            }
            s->reserved -= fp;
            s->running--;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    if (ctx) gt_ctx_destroy(ctx);
    if (st) (void)hipStreamDestroy(st);
}

} // namespace

extern "C" {

int gt_server_create(gt_server **out, int device, const gt_server_config *cfg) {
    if (!out) return gt::set_error(GT_EINVAL, "null server pointer");
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return gt::set_error(GT_EDEVICE, "server: device unavailable");
    }
    gt_server *s = new gt_server();
    s->device = device;
    s->density = gt::density_table_new();
    if (cfg) s->cfg = *cfg;
    if (s->cfg.slots <= 0) s->cfg.slots = 4;
    if (s->cfg.slots > 16) s->cfg.slots = 16;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 16;
    if (s->cfg.batch_pixels <= 0) s->cfg.batch_pixels = 1 << 22;
    if (s->cfg.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            fr = (size_t)16 << 30;
        }
        s->cfg.hbm_budget = fr / 2;
    }
    if (s->cfg.warm_side > 0) {
        // dense content: independent uniform noise per channel (xorshift)
        const size_t n = (size_t)s->cfg.warm_side * s->cfg.warm_side;
        s->warm.resize(3 * n);
        uint64_t x = 0x9E3779B97F4A7C15ull;
        for (size_t i = 0; i < 3 * n; i++) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            s->warm[i] = (float)(x >> 40) * (1.0f / 16777216.0f);
        }
    }
    std::mutex init_mu;
    std::condition_variable init_cv;
    int left = s->cfg.slots;
    for (int i = 0; i < s->cfg.slots; i++) s->workers.emplace_back(worker_main, s, &init_mu, &left, &init_cv);
    {
        std::unique_lock<std::mutex> lk(init_mu);
        init_cv.wait(lk, [&] { return left == 0; });
    }
    s->warm.clear();
    s->warm.shrink_to_fit();
    if (s->init_rc != GT_OK) {
        const int rc = s->init_rc;
        const std::string m = s->init_err;
        gt_server_destroy(s);
        return gt::set_error(rc, m.c_str());
    }
    *out = s;
    return GT_OK;
}

void gt_server_destroy(gt_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &kv : s->reqs) delete kv.second;   // every queued request is also in reqs
    gt::density_table_delete(s->density);         // after the workers' contexts are gone
    delete s;
}

int gt_server_submit(gt_server *s, const gt_request *r, long long *ticket) {
    if (!s || !r || !ticket) return gt::set_error(GT_EINVAL, "server: null argument");
    if (r->W <= 0 || r->H <= 0 || (r->C != 1 && r->C != 3) || !r->im || !r->out || !(r->sigma_s > 0) ||
        !(r->sigma_r > 0) || r->row_pitch < r->W || r->plane_pitch < (long long)r->row_pitch * (r->H - 1) + r->W ||
        (int64_t)r->W * r->H >= (int64_t)1 << 31 || !(r->deadline_ms >= 0))
        return gt::set_error(GT_EINVAL, "server: bad request (sizes, C in {1,3}, sigmas > 0, pitches, deadline >= 0)");
    Req *q = new Req();
    q->r = *r;
    q->fp = footprint(*r);
    if (q->fp > s->cfg.hbm_budget) {
        delete q;
        return gt::set_error(GT_ENOMEM, "server: request footprint exceeds the HBM budget");
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        q->ticket = s->next_ticket++;
        q->t_submit = now_ms();
        s->reqs[q->ticket] = q;
        s->queue.push_back(q);
        *ticket = q->ticket;
    }
    s->cv_work.notify_one();
    return GT_OK;
}

int gt_server_wait(gt_server *s, long long ticket, gt_request_result *res) {
    if (!s) return gt::set_error(GT_EINVAL, "server: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return gt::set_error(GT_EINVAL, "server: unknown ticket");
    Req *q = it->second;
    if (q->r.deadline_ms > 0 && !q->taken && !q->done) {
        // expire it here at its deadline if every worker is still busy then
        using namespace std::chrono;
        const steady_clock::time_point at(
            duration_cast<steady_clock::duration>(duration<double, std::milli>(q->t_submit + q->r.deadline_ms)));
        if (!s->cv_done.wait_until(lk, at, [&] { return q->done || q->taken; }) && !q->taken && !q->done) {
            s->queue.erase(std::find(s->queue.begin(), s->queue.end(), q));
            finish_unrun(q, GT_ETIMEDOUT, "server: request not started before its deadline", now_ms());
            s->cv_work.notify_all();   // the queue head may have changed
        }
    }
    s->cv_done.wait(lk, [&] { return q->done; });
    if (res) {
        res->status = q->status;
        res->batch = q->batch;
        res->latency_ms = q->t_done - q->t_submit;
        res->queue_ms = q->t_start - q->t_submit;
        res->footprint = q->fp;
    }
    const int rc = q->status;
    const std::string m = q->err;
    s->reqs.erase(it);
    delete q;
    lk.unlock();
    if (rc != GT_OK) return gt::set_error(rc, m.c_str());
    return GT_OK;
}

int gt_server_cancel(gt_server *s, long long ticket) {
    if (!s) return gt::set_error(GT_EINVAL, "server: null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return gt::set_error(GT_EINVAL, "server: unknown ticket");
        Req *q = it->second;
        if (q->taken || q->done) return gt::set_error(GT_ESTATE, "server: request already running or done");
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), q));
        finish_unrun(q, GT_ECANCELED, "server: request cancelled", now_ms());
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();   // the queue head may have changed
    return GT_OK;
}

int gt_server_replay(gt_server *s, int count, const gt_request *reqs, const double *arrival_ms,
                     gt_request_result *results) {
    if (!s || count < 0 || (count && (!reqs || !arrival_ms || !results)))
        return gt::set_error(GT_EINVAL, "server: bad replay arguments");
    std::vector<long long> tickets(count, 0);
    const auto t0 = std::chrono::steady_clock::now();
    int first_rc = GT_OK;
    for (int i = 0; i < count; i++) {
        std::this_thread::sleep_until(t0 + std::chrono::duration<double, std::milli>(arrival_ms[i]));
        const int rc = gt_server_submit(s, &reqs[i], &tickets[i]);
        if (rc != GT_OK) {
            tickets[i] = 0;
            results[i] = gt_request_result{};
            results[i].status = rc;
            if (first_rc == GT_OK) first_rc = rc;
        }
    }
    std::string m;
    for (int i = 0; i < count; i++) {
        if (!tickets[i]) continue;
        const int rc = gt_server_wait(s, tickets[i], &results[i]);
        if (rc != GT_OK && first_rc == GT_OK) {
            first_rc = rc;
            m = gt_last_error();
        }
    }
    if (first_rc != GT_OK && !m.empty()) gt::set_error(first_rc, m.c_str());
    return first_rc;
}

} // extern "C"
