// Autotuner: times the lattice pipeline over a grid of tunings (pixel-tile
// width, hash-table load factor, waves per CU of the tiled and neighbour
// kernels) on the caller's own inputs and returns the fastest.  Built only on
// the public C ABI; trial outputs go to a scratch buffer, so inputs that alias
// outputs are never overwritten.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <vector>

#include "gt_internal.h"
#include "../../include/gtlattice.h"

namespace {

const int kTileW[] = {16, 32, 64};
const float kLoad[] = {0.25f, 0.35f, 0.5f};
const int kWaves[] = {0, 12, 20};

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Median wall time (ms) of `reps` calls after one untimed call.  Each call is
// synchronous (it returns when its results are complete).
int time_calls(const std::function<int()> &call, int reps, float *ms) {
    int rc = call();
    if (rc != GT_OK) return rc;
    std::vector<double> t;
    for (int i = 0; i < reps; i++) {
        const double t0 = now_ms();
        rc = call();
        if (rc != GT_OK) return rc;
        t.push_back(now_ms() - t0);
    }
    std::sort(t.begin(), t.end());
    *ms = (float)t[t.size() / 2];
    return GT_OK;
}

int search(const std::function<int(const gt_tuning &)> &run, int reps, gt_tuning *best, gt_tune_report *rep) {
    gt_tuning def{};
    float def_ms = 0.f;
    int rc = time_calls([&] { return run(def); }, reps, &def_ms);
    if (rc != GT_OK) return rc;
    gt_tuning b = def;
    float b_ms = def_ms;
    int trials = 1;
    for (int tw : kTileW)
        for (float lf : kLoad)
            for (int w : kWaves) {
                gt_tuning t{};
                t.tile_w = tw;
                t.load_factor = lf;
                t.waves_per_cu = w;
                float ms = 0.f;
                rc = time_calls([&] { return run(t); }, reps, &ms);
                if (rc != GT_OK) return rc;
                trials++;
                if (ms < b_ms) {
                    b_ms = ms;
                    b = t;
                }
            }
    // then the neighbour-table form (sorted ids + merge, or hash lookups), on the best of the rest
    for (int nb : {1, 2}) {
        gt_tuning t = b;
        t.neighbours = nb;
        float ms = 0.f;
        rc = time_calls([&] { return run(t); }, reps, &ms);
        if (rc != GT_OK) return rc;
        trials++;
        if (ms < b_ms) {
            b_ms = ms;
            b = t;
        }
    }
    if (best) *best = b;
    if (rep) {
        rep->trials = trials;
        rep->best_ms = b_ms;
        rep->default_ms = def_ms;
    }
    return GT_OK;
}

struct Scratch {
    float *p = nullptr;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
};

} // namespace

extern "C" int gt_autotune_joint_bilateral(gt_ctx *ctx, int W, int H, int F, int C, const float *im, int Cr,
                                           const float *ref, float fw, float fh, float ff, float color_sigma,
                                           gt_storage storage, int reps, gt_tuning *best, gt_tune_report *rep) {
    if (!ctx || W <= 0 || H <= 0 || F <= 0 || C <= 0) return gt::set_error(GT_EINVAL, "autotune: bad sizes");
    Scratch out;
    if (gt::alloc_should_fail() || hipMalloc((void **)&out.p, (size_t)W * H * F * C * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return gt::set_error(GT_ENOMEM, "autotune: no device memory for the trial output");
    }
    auto run = [&](const gt_tuning &t) {
        return gt_joint_bilateral(ctx, W, H, F, C, im, Cr, ref, fw, fh, ff, color_sigma, out.p, storage, &t, nullptr);
    };
    return search(run, reps < 1 ? 3 : reps, best, rep);
}

extern "C" int gt_autotune_joint_bilateral_batch(gt_ctx *ctx, int count, const gt_bilateral_image *images, int C,
                                                 int Cr, gt_storage storage, int reps, gt_tuning *best,
                                                 gt_tune_report *rep) {
    if (!ctx || count <= 0 || !images || C <= 0) return gt::set_error(GT_EINVAL, "autotune: bad batch");
    size_t total = 0;
    for (int i = 0; i < count; i++) total += (size_t)images[i].W * images[i].H * C;
    Scratch out;
    if (gt::alloc_should_fail() || hipMalloc((void **)&out.p, total * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return gt::set_error(GT_ENOMEM, "autotune: no device memory for the trial outputs");
    }
    std::vector<gt_bilateral_image> imgs(images, images + count);
    size_t off = 0;
    for (auto &b : imgs) {
        b.out = out.p + off;
        off += (size_t)b.W * b.H * C;
    }
    auto run = [&](const gt_tuning &t) {
        return gt_joint_bilateral_batch(ctx, count, imgs.data(), C, Cr, storage, &t, nullptr);
    };
    return search(run, reps < 1 ? 3 : reps, best, rep);
}
