"""The formats either side of the path on the GPU: .tmp files straight into
and out of device memory (samples converted on the GPU, bit-identical to the
host reader/writer and so to the reference's files), and the -send /
-receive front end running the bilateral filter through the serving loop,
driven by a socket client and by the unmodified reference CLI."""
import os
import socket
import subprocess

import numpy as np
import pytest

from imagestack_amd import gtlattice as G
from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TMPD = os.path.join(HERE, "golden", "tmp")
CLI = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ImageStack")


def raw(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("t", list(G.TMP_TYPES))
def test_tmp_device_roundtrip_matches_reference(gpu_ctx, tmp_path, t):
    print(f"\n[case] .tmp {t} through device memory", flush=True)
    d = gpu_ctx.tmp_load(os.path.join(TMPD, f"as_{t}.tmp"))
    want = G.tmp_read(os.path.join(TMPD, f"back_{t}.tmp"))
    assert np.array_equal(d.cpu().numpy().view(np.uint32), want.view(np.uint32))
    src = gpu_ctx.tmp_load(os.path.join(TMPD, "source.tmp"))
    out = str(tmp_path / "out.tmp")
    gpu_ctx.tmp_save(out, src, t)
    assert raw(out) == raw(os.path.join(TMPD, f"as_{t}.tmp"))


@pytest.mark.parametrize("t", ["float32", "uint8", "float64"])
def test_tmp_large_pipelined(gpu_ctx, tmp_path, t):
    """Files larger than the 16 MiB staging chunks (several chunks in flight)."""
    print(f"\n[case] .tmp 2048x2048x3 {t} load/save", flush=True)
    im = (S.uniform_image(2048, 2048, 3, seed=5) * 255.0).astype(np.float32)
    p = str(tmp_path / "big.tmp")
    G.tmp_write(p, im, t)
    d = gpu_ctx.tmp_load(p)
    assert np.array_equal(d.cpu().numpy(), G.tmp_read(p))
    q = str(tmp_path / "big2.tmp")
    gpu_ctx.tmp_save(q, d, t)
    assert raw(q) == raw(p)


def test_net_bilateral_socket_client(oracle_lib):
    print("\n[case] -send/-receive front end, bilateral, same-connection replies", flush=True)
    srv = G.NetServer(op=G.GT_NET_BILATERAL, sigma_s=6.0, sigma_r=0.15, slots=2)
    try:
        cases = [S.rect_image(96, 64, 3, K=8, noise=0.05, seed=61), S.rect_image(80, 50, 1, K=6, noise=0.05, seed=62),
                 S.rect_image(40, 30, 2, K=4, noise=0.05, seed=63, F=3)]
        for im in cases:
            with socket.create_connection(("127.0.0.1", srv.port), timeout=60) as s:
                G.send_image(s, im)
                out = G.recv_image(s)
            o64 = oracle_lib.joint_bilateral(im, im, 6.0, 6.0, 0.0, 0.15, mode=oracle_lib.F64_LAT)
            assert out.shape == im.shape
            assert rel_err(out, o64) < 2 * FP32_REL_TOL
        st = srv.stats()
        assert st["replied"] == 3 and st["failed"] == 0 and st["p99_ms"] > 0
    finally:
        srv.close()


def test_net_bilateral_reference_cli(tmp_path, oracle_lib):
    """ImageStack -load a.tmp -send 127.0.0.1 P -receive R -save b.tmp with the
    unmodified reference CLI: b is a's bilateral (sigma_s 8, sigma_r 0.1)."""
    if not os.path.exists(CLI):
        pytest.skip("reference CLI not built")
    print("\n[case] reference CLI -send/-receive through the GPU front end", flush=True)
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        reply = s.getsockname()[1]
    srv = G.NetServer(op=G.GT_NET_BILATERAL, reply_port=reply, sigma_s=8.0, sigma_r=0.1, slots=1)
    try:
        im = S.rect_image(128, 96, 3, K=10, noise=0.05, seed=64)
        src, dst = str(tmp_path / "a.tmp"), str(tmp_path / "b.tmp")
        G.tmp_write(src, im)
        r = subprocess.run([CLI, "-load", src, "-send", "127.0.0.1", str(srv.port), "-receive", str(reply),
                            "-save", dst], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        o64 = oracle_lib.joint_bilateral(im, im, 8.0, 8.0, 0.0, 0.1, mode=oracle_lib.F64_LAT)
        assert rel_err(G.tmp_read(dst), o64) < 2 * FP32_REL_TOL
    finally:
        srv.close()
