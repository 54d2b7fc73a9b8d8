"""Named lattice workloads, run K times on one GPU, for profiling a single
configuration under rocprofv3 (the benchmark runs several in one process).

    python -m imagestack_amd.workloads jb2048   [--steps K] [--warmup W] [--eager]
    python -m imagestack_amd.workloads highdim  [--storage f32|bf16|f16]
    python -m imagestack_amd.workloads batch
    python -m imagestack_amd.workloads nlmeans

jb2048  : joint bilateral 2048x2048 RGB by an RGB guide, sigma_s 16, sigma_r
          0.1 (d = 5, vd = 4).
highdim : Gauss transform at 1024x1024 with d = 16 positions (14 N(0,1)
          features + x/12, y/12) and 4 values U[0,1], unnormalised.
batch   : 64 RGB images, sides log-uniform in [256, 1536], per-image sigmas,
          as one batched lattice.
nlmeans : non-local means on a noisy 2048x2048 RGB image (7x7 patches, 6 PCA
          dimensions, sigma_s 16, patch sigma 0.4; d = 8 lattice).

The pipeline runs as a replayed graph unless --eager (launch by launch, with
per-stage HIP events).  --tune takes gt_tuning fields as JSON, e.g.
'{"load_factor": 0.5}'.  Prints one JSON line with the mean wall time per call.
"""
import argparse
import json
import os
import sys
import time

if __package__ in (None, ""):   # run as a file: the package's parent on the path
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["jb2048", "highdim", "batch", "nlmeans"])
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--storage", default="f32", choices=["f32", "bf16", "f16"])
    ap.add_argument("--eager", action="store_true")
    ap.add_argument("--tune", default=None, help="gt_tuning fields as JSON")
    a = ap.parse_args()
    tune = json.loads(a.tune) if a.tune else None

    import numpy as np
    import torch

    from imagestack_amd import gtlattice, synth

    if not torch.cuda.is_available():
        raise SystemExit("a GPU is needed")
    ctx = gtlattice.Context(0, profiling=1 if a.eager else 0)
    if a.workload == "jb2048":
        im, guide, ss, sr = synth.joint_bilateral_2048()
        d_im, d_g = torch.from_numpy(im).cuda(), torch.from_numpy(guide).cuda()
        out = torch.empty_like(d_im)
        npx = im.shape[-1] * im.shape[-2]

        def call():
            return ctx.joint_bilateral(d_im, d_g, ss, ss, 0.0, sr, out=out, storage=a.storage, tune=tune,
                                       stats=True)[1]
    elif a.workload == "highdim":
        side = 1024
        pos = synth.highdim_positions(side, side, nfeat=14, sigma_s=12.0, seed=3)
        vals = synth.uniform_image(side, side, 4, seed=33)
        d = pos.shape[0]
        dp = torch.from_numpy(pos.reshape(d, -1)).cuda()
        dv = torch.from_numpy(vals.reshape(4, -1)).cuda()
        inv = np.ones(d, np.float32)
        npx = side * side

        def call():
            return ctx.gauss_transform(dp, dv, inv, row_w=side, storage=a.storage, tune=tune, stats=True)[1]
    elif a.workload == "batch":
        imgs = synth.batch_images()
        dev = [torch.from_numpy(im[:, None]).cuda() for im, _, _ in imgs]
        items = [(t, t, torch.empty_like(t), ss, sr) for t, (_, ss, sr) in zip(dev, imgs)]
        npx = sum(im.shape[1] * im.shape[2] for im, _, _ in imgs)

        def call():
            return ctx.joint_bilateral_batch(items, storage=a.storage, tune=tune, stats=True)
    else:
        from imagestack_amd import nlmeans as NL
        im = synth.noisy_rect_image(2048, 2048, 3, K=64, amp=0.2, seed=46)
        filt = NL.patch_pca(im, 1.0, 6, seed=0)
        d_im = torch.from_numpy(im).cuda()
        d_f = torch.from_numpy(filt).cuda()
        npx = 2048 * 2048

        def call():
            return ctx.nlmeans(d_im, d_f, 16.0, 0.4, storage=a.storage, tune=tune, stats=True)[1]

    for _ in range(a.warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st = call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    ctx.close()
    print(json.dumps({"workload": a.workload, "storage": a.storage, "eager": a.eager, "tune": tune, "steps": a.steps,
                      "ms_per_call": round(ms, 4), "mpix_s": round(npx / ms / 1e3, 2), "M": st["M"],
                      "capacity": st["capacity"],
                      "stages_ms": {k: round(st[k], 4) for k in ("ms_splat", "ms_sort", "ms_nbr", "ms_blur",
                                                                 "ms_slice", "ms_total")} if a.eager else None}),
          flush=True)


if __name__ == "__main__":
    main()
