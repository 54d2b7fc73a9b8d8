"""Failure paths of the request loop and of a context.

  * Cancellation: a request still queued completes with GT_ECANCELED and its
    output is not written; one already done answers GT_ESTATE; the other
    requests still give the output of the same request filtered alone.
  * Deadlines: a request that no worker takes within its deadline completes
    with GT_ETIMEDOUT, unrun, whether its waiter or a worker notices first; a
    request taken in time runs normally.
  * Allocation failure (gt_debug_fail_allocations): a failure at any one of
    the allocations of a call, a server request or a context's creation is
    returned as GT_ENOMEM, and the same context or server then gives the
    output a fresh one gives, bit for bit.
"""
import numpy as np
import pytest

from imagestack_amd import gtlattice as G
from imagestack_amd import synth as S
from tests.util import rel_err

pytestmark = pytest.mark.gpu

SENTINEL = np.float32(-7.0)
MIX_TOL = 1e-6


def case(name):
    print(f"\n[case] {name}", flush=True)


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def alone(im, ss, sr):
    ctx = G.Context(0)
    try:
        return ctx.joint_bilateral_host(im[:, None], im[:, None], ss, ss, 0.0, sr)[:, 0]
    finally:
        ctx.close()


def small(seed, W=160, H=120, C=3):
    return np.ascontiguousarray(S.rect_image(W, H, C, K=8, noise=0.05, seed=seed)[:, 0])


@pytest.fixture(scope="module")
def long_request():
    """A request that keeps the one worker busy for well over the few
    milliseconds the tests need: 2048^2 RGB from pageable memory, sigma_s 4."""
    im = np.ascontiguousarray(S.uniform_image(2048, 2048, 3, seed=5)[:, 0])
    return im, np.zeros_like(im)


@pytest.fixture(autouse=True)
def no_injection():
    G.fail_allocations(0)
    yield
    G.fail_allocations(0)


def test_cancel_queued_request(long_request):
    case("cancel: one worker busy on 2048^2 RGB, three 160x120 requests queued")
    big, big_out = long_request
    ims = [small(400 + k) for k in range(3)]
    outs = [np.full_like(im, SENTINEL) for im in ims]
    srv = G.Server(0, slots=1, max_batch=1)
    try:
        t0 = srv.submit(G.make_request(big, big_out, 4.0, 0.1))
        ts = [srv.submit(G.make_request(im, o, 8.0, 0.1)) for im, o in zip(ims, outs)]
        srv.cancel(ts[1])
        with pytest.raises(G.GTError) as e:
            srv.cancel(ts[1])        # already done (cancelled)
        assert e.value.code == G.GT_ESTATE
        r2 = srv.wait(ts[2])
        assert r2["status"] == 0
        with pytest.raises(G.GTError) as e:
            srv.cancel(t0)           # ran to the end, not yet collected
        assert e.value.code == G.GT_ESTATE
        r1 = srv.wait(ts[1], check=False)
        assert r1["status"] == G.GT_ECANCELED
        with pytest.raises(G.GTError) as e:
            srv.wait(ts[1])          # collected
        assert e.value.code == G.GT_EINVAL
        with pytest.raises(G.GTError) as e:
            srv.cancel(ts[1])        # unknown now
        assert e.value.code == G.GT_EINVAL
        assert srv.wait(t0)["status"] == 0 and srv.wait(ts[0])["status"] == 0
        # the server keeps serving
        extra = small(410)
        extra_out = np.full_like(extra, SENTINEL)
        assert srv.wait(srv.submit(G.make_request(extra, extra_out, 8.0, 0.1)))["status"] == 0
    finally:
        srv.close()
    assert np.all(outs[1] == SENTINEL), "a cancelled request's output was written"
    for k in (0, 2):
        assert rel_err(outs[k], alone(ims[k], 8.0, 0.1)) < MIX_TOL, k
    assert rel_err(extra_out, alone(extra, 8.0, 0.1)) < MIX_TOL
    assert 0.3 < float(big_out.mean()) < 0.7   # the long request ran


def test_deadlines(long_request):
    case("deadlines: expired in the waiter, expired by the worker, taken in time")
    big, big_out = long_request
    ims = [small(500 + k) for k in range(4)]
    outs = [np.full_like(im, SENTINEL) for im in ims]
    srv = G.Server(0, slots=1, max_batch=1)
    try:
        with pytest.raises(G.GTError) as e:
            srv.submit(G.make_request(ims[0], outs[0], 8.0, 0.1, deadline_ms=-1.0))
        assert e.value.code == G.GT_EINVAL
        # the waiter notices: it waits on the 1 ms request while the worker is busy
        t0 = srv.submit(G.make_request(big, big_out, 4.0, 0.1))
        ta = srv.submit(G.make_request(ims[0], outs[0], 8.0, 0.1, deadline_ms=1.0))
        tb = srv.submit(G.make_request(ims[1], outs[1], 8.0, 0.1, deadline_ms=60000.0))
        ra = srv.wait(ta, check=False)
        assert ra["status"] == G.GT_ETIMEDOUT
        assert ra["latency_ms"] >= 0.99 and ra["batch"] == 0
        assert srv.wait(t0)["status"] == 0
        assert srv.wait(tb)["status"] == 0
        # the worker notices: nobody waits on the 1 ms request until later
        t0 = srv.submit(G.make_request(big, big_out, 4.0, 0.1))
        tc = srv.submit(G.make_request(ims[2], outs[2], 8.0, 0.1, deadline_ms=1.0))
        td = srv.submit(G.make_request(ims[3], outs[3], 8.0, 0.1))
        assert srv.wait(td)["status"] == 0
        rc = srv.wait(tc, check=False)
        assert rc["status"] == G.GT_ETIMEDOUT and rc["batch"] == 0
        assert srv.wait(t0)["status"] == 0
    finally:
        srv.close()
    for k in (0, 2):
        assert np.all(outs[k] == SENTINEL), f"an expired request's output was written ({k})"
    for k in (1, 3):
        assert rel_err(outs[k], alone(ims[k], 8.0, 0.1)) < MIX_TOL, k


def jb(ctx, im, guide, ss, sr):
    import torch
    a = torch.from_numpy(im).cuda()
    g = torch.from_numpy(guide).cuda()
    return ctx.joint_bilateral(a, g, ss, ss, 0.0, sr).cpu().numpy()


def test_context_allocation_failure_at_every_point():
    """A context that ran a 128x96 filter meets a failed allocation at the
    k-th allocation of a 512x384 filter (which grows every buffer), for every
    k: that call returns GT_ENOMEM, and the next calls on the same context,
    at both sizes, give a fresh context's bits."""
    case("context: injected failure at each allocation of a growing call")
    g1 = S.rect_image(128, 96, 3, K=10, noise=0.05, seed=31)
    i1 = S.uniform_image(128, 96, 3, seed=32)
    g2 = S.rect_image(512, 384, 3, K=20, noise=0.05, seed=33)
    i2 = S.uniform_image(512, 384, 3, seed=34)
    fresh = G.Context(0)
    try:
        want1 = jb(fresh, i1, g1, 6.0, 0.1)
    finally:
        fresh.close()
    fresh = G.Context(0)
    try:
        want2 = jb(fresh, i2, g2, 6.0, 0.1)
    finally:
        fresh.close()
    failed_at = []
    for k in range(64):
        ctx = G.Context(0)
        try:
            assert same_bits(jb(ctx, i1, g1, 6.0, 0.1), want1)
            G.fail_allocations(1, after=k)
            try:
                jb(ctx, i2, g2, 6.0, 0.1)
                hit = False
            except G.GTError as e:
                assert e.code == G.GT_ENOMEM, e
                hit = True
            G.fail_allocations(0)
            assert same_bits(jb(ctx, i2, g2, 6.0, 0.1), want2), k
            assert same_bits(jb(ctx, i1, g1, 6.0, 0.1), want1), k
            assert same_bits(jb(ctx, i2, g2, 6.0, 0.1), want2), k
        finally:
            G.fail_allocations(0)
            ctx.close()
        if not hit:
            break
        failed_at.append(k)
    print("failures injected at allocations", failed_at, flush=True)
    assert len(failed_at) >= 5 and not hit


def test_context_creation_allocation_failure():
    case("context creation: injected pinned-allocation failure")
    G.fail_allocations(1)
    with pytest.raises(G.GTError) as e:
        G.Context(0)
    assert e.value.code == G.GT_ENOMEM
    G.fail_allocations(0)
    ctx = G.Context(0)
    ctx.close()


def test_server_allocation_failure_at_every_point():
    """A failed allocation at the k-th allocation a server request makes (the
    worker's plane arena, its pinned staging, its context's buffers) returns
    GT_ENOMEM for that request only; the next request on the same server is
    right."""
    case("server: injected failure at each allocation of the first request")
    im = small(600, 200, 150)
    want = alone(im, 8.0, 0.1)
    failed_at = []
    for k in range(64):
        srv = G.Server(0, slots=1, max_batch=1)
        try:
            out = np.full_like(im, SENTINEL)
            G.fail_allocations(1, after=k)
            r = srv.wait(srv.submit(G.make_request(im, out, 8.0, 0.1)), check=False)
            G.fail_allocations(0)
            hit = r["status"] != 0
            if hit:
                assert r["status"] == G.GT_ENOMEM, r
            else:
                assert rel_err(out, want) < MIX_TOL
            out2 = np.full_like(im, SENTINEL)
            assert srv.wait(srv.submit(G.make_request(im, out2, 8.0, 0.1)))["status"] == 0
            assert rel_err(out2, want) < MIX_TOL, k
        finally:
            G.fail_allocations(0)
            srv.close()
        if not hit:
            break
        failed_at.append(k)
    print("failures injected at allocations", failed_at, flush=True)
    assert len(failed_at) >= 3 and not hit
