"""The request loop (gt_server_*) and the autotuner.

Serving must not change results: every request of a mixed, concurrently
submitted set (grey and RGB, uneven sizes, batched together by the server)
matches the same request filtered alone, within 1e-6 relative (the only
difference a batch makes is its fixed-point scale), and the FP64 oracle within
the 1e-5 gate.  The autotuner's choices only move work around: every tuning
gives the bit-identical output, because the splat sums are exact."""
import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-6


def make_requests(seed=0, count=14):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        C = 1 if k % 5 in (0, 2, 4) else 3
        W = int(rng.integers(40, 260))
        H = int(rng.integers(30, 220))
        im = S.rect_image(W, H, C, K=8, noise=0.05, seed=3000 + seed * 100 + k)[:, 0]
        ss = float(rng.choice([4.0, 16.0, 32.0]))
        sr = float(rng.choice([0.05, 0.2]))
        reqs.append((np.ascontiguousarray(im), ss, sr))
    return reqs


def alone(im, ss, sr, storage="f32"):
    from imagestack_amd import gtlattice
    ctx = gtlattice.Context(0)
    try:
        out = ctx.joint_bilateral_host(im[:, None], im[:, None], ss, ss, 0.0, sr, storage=storage)
    finally:
        ctx.close()
    return out[:, 0]


def test_mix_equals_alone_and_oracle(oracle_lib):
    from imagestack_amd import gtlattice
    print("\n[case] 14 mixed grey/RGB requests, server slots=3 max_batch=6", flush=True)
    reqs = make_requests(1)
    srv = gtlattice.Server(0, slots=3, max_batch=6)
    try:
        outs = [np.zeros_like(im) for im, _, _ in reqs]
        handles = [gtlattice.make_request(im, o, ss, sr) for (im, ss, sr), o in zip(reqs, outs)]
        tickets = [srv.submit(h) for h in handles]
        results = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    assert all(r["status"] == 0 for r in results)
    assert max(r["batch"] for r in results) > 1, "expected the server to batch waiting requests"
    for k, ((im, ss, sr), got) in enumerate(zip(reqs, outs)):
        ref = alone(im, ss, sr)
        assert rel_err(got, ref) < MIX_TOL, k
        if k < 6:
            o64 = oracle_lib.joint_bilateral(im[:, None], im[:, None], ss, ss, 0.0, sr, mode=oracle_lib.F64_LAT)
            assert rel_err(got, o64[:, 0]) < FP32_REL_TOL, k


def test_admission_budget_and_strided_input():
    """A budget that admits about one request at a time still completes every
    request correctly; a request beyond the budget is refused at submit; a
    request may read a crop of a larger planar host image."""
    from imagestack_amd import gtlattice
    print("\n[case] tight HBM budget, crop requests, oversize refusal", flush=True)
    big = S.rect_image(300, 260, 3, K=12, noise=0.05, seed=77)[:, 0]
    crops = [(10, 20, 200, 150, 8.0, 0.1), (0, 0, 256, 256, 16.0, 0.2), (50, 40, 120, 90, 4.0, 0.05)]
    probe = gtlattice.make_request(big, np.zeros((3, 256, 256), np.float32), 16.0, 0.2)
    srv = gtlattice.Server(0, slots=2, max_batch=4, hbm_budget=int(1.2 * _footprint_of(probe)))
    try:
        outs, tickets = [], []
        for (x, y, W, H, ss, sr) in crops:
            o = np.zeros((3, H, W), np.float32)
            addr = big.ctypes.data + 4 * (y * 300 + x)
            req = gtlattice.make_request(addr, o, ss, sr, row_pitch=300, plane_pitch=300 * 260, W=W, H=H)
            outs.append(o)
            tickets.append(srv.submit(req))
        res = [srv.wait(t) for t in tickets]
        huge = np.zeros((3, 4096, 4096), np.float32)
        with pytest.raises(gtlattice.GTError):
            srv.submit(gtlattice.make_request(huge, huge, 8.0, 0.1))
    finally:
        srv.close()
    assert all(r["status"] == 0 for r in res)
    for (x, y, W, H, ss, sr), got in zip(crops, outs):
        dense = np.ascontiguousarray(big[:, y:y + H, x:x + W])
        assert rel_err(got, alone(dense, ss, sr)) < MIX_TOL


def _footprint_of(req):
    """The server's footprint estimate of one request, read back from a
    one-request run on an unconstrained server."""
    from imagestack_amd import gtlattice
    im = np.zeros((req.C, req.H, req.W), np.float32)
    out = np.zeros_like(im)
    srv = gtlattice.Server(0, slots=1)
    try:
        t = srv.submit(gtlattice.make_request(im, out, req.sigma_s, req.sigma_r))
        return srv.wait(t)["footprint"]
    finally:
        srv.close()


def test_bf16_requests_match_alone():
    from imagestack_amd import gtlattice
    print("\n[case] bf16-stored requests through the server", flush=True)
    reqs = make_requests(2, count=6)
    srv = gtlattice.Server(0, slots=2, max_batch=4)
    try:
        outs = [np.zeros_like(im) for im, _, _ in reqs]
        tickets = [srv.submit(gtlattice.make_request(im, o, ss, sr, storage="bf16"))
                   for (im, ss, sr), o in zip(reqs, outs)]
        for t in tickets:
            assert srv.wait(t)["status"] == 0
    finally:
        srv.close()
    for (im, ss, sr), got in zip(reqs, outs):
        ref = alone(im, ss, sr, storage="bf16")
        # bf16 rounding of the blurred values can flip on the last fixed-point
        # bit of a batch; the error stays far inside the bf16 storage bar.
        assert np.linalg.norm(got - ref) <= 1e-3 * np.linalg.norm(ref)


def test_replay_reports_latency():
    from imagestack_amd import gtlattice
    print("\n[case] open-loop replay of 10 requests at 200 QPS", flush=True)
    reqs = make_requests(3, count=10)
    outs = [np.zeros_like(im) for im, _, _ in reqs]
    handles = [gtlattice.make_request(im, o, ss, sr) for (im, ss, sr), o in zip(reqs, outs)]
    arrivals = [5.0 * k for k in range(len(handles))]
    srv = gtlattice.Server(0, slots=2)
    try:
        res = srv.replay(handles, arrivals)
    finally:
        srv.close()
    assert all(r["status"] == 0 and r["latency_ms"] > 0 and r["queue_ms"] >= 0 for r in res)
    for (im, ss, sr), got in zip(reqs, outs):
        assert rel_err(got, alone(im, ss, sr)) < MIX_TOL


def test_autotune_choices_give_identical_output(gpu_ctx):
    import torch
    print("\n[case] autotune 256x192 RGB; outputs identical across tunings", flush=True)
    im = S.rect_image(256, 192, 3, K=16, noise=0.05, seed=55)
    d_im = torch.from_numpy(im).cuda()
    best, rep = gpu_ctx.autotune_joint_bilateral(d_im, d_im, 8.0, 8.0, 0.0, 0.1, reps=2)
    assert rep["trials"] == 30 and rep["best_ms"] <= rep["default_ms"]   # 27 grid points, defaults, 2 forms
    assert best["tile_w"] in (0, 16, 32, 64)
    outs = []
    for tune in (None, best, {"tile_w": 16, "load_factor": 0.5, "waves_per_cu": 12},
                 {"tile_w": 64, "load_factor": 0.25, "waves_per_cu": 20}, {"neighbours": 1}, {"neighbours": 2}):
        outs.append(gpu_ctx.joint_bilateral(d_im, d_im, 8.0, 8.0, 0.0, 0.1, tune=tune).cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
