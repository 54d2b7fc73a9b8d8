"""The .tmp file format and the -send / -receive wire format, host side (no
GPU): libgtlattice.so's gt_tmp_write / gt_tmp_read against files the
reference itself wrote and read (tests/golden/tmp: a float image saved as
every sample type by the reference CLI, and each loaded back), byte for byte;
the old-style header; and the TCP front end in echo mode driven by the
unmodified reference CLI (-send / -receive) and by a socket client."""
import os
import socket
import subprocess
import time

import numpy as np
import pytest

from imagestack_amd import gtlattice as G

HERE = os.path.dirname(os.path.abspath(__file__))
TMPD = os.path.join(HERE, "golden", "tmp")
CLI = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ImageStack")
TYPES = list(G.TMP_TYPES)


def raw(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("t", TYPES)
def test_write_matches_reference_bytes(tmp_path, t):
    src = G.tmp_read(os.path.join(TMPD, "source.tmp"))
    out = str(tmp_path / f"{t}.tmp")
    G.tmp_write(out, src, t)
    assert raw(out) == raw(os.path.join(TMPD, f"as_{t}.tmp"))
    assert G.tmp_info(out) == ((7, 5, 2, 3), G.TMP_TYPES[t])


@pytest.mark.parametrize("t", TYPES)
def test_read_matches_reference_load(t):
    got = G.tmp_read(os.path.join(TMPD, f"as_{t}.tmp"))
    want = G.tmp_read(os.path.join(TMPD, f"back_{t}.tmp"))
    assert got.dtype == np.float32 and np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_conversion_semantics():
    """What the fixtures hold, stated independently: truncation toward zero,
    low bits kept for negatives in unsigned types, exact widening."""
    src = G.tmp_read(os.path.join(TMPD, "source.tmp")).reshape(-1)
    assert list(src[:6]) == [0.0, 0.0, 0.5, -0.5, np.float32(99.99), np.float32(-99.99)]
    u8 = np.fromfile(os.path.join(TMPD, "as_uint8.tmp"), np.uint8, offset=20)
    assert list(u8[:6]) == [0, 0, 0, 0, 99, 157]        # -99 -> 157 (mod 256)
    i16 = np.fromfile(os.path.join(TMPD, "as_int16.tmp"), np.int16, offset=20)
    np.testing.assert_array_equal(i16, np.trunc(src).astype(np.int16))
    f64 = np.fromfile(os.path.join(TMPD, "as_float64.tmp"), np.float64, offset=20)
    np.testing.assert_array_equal(f64, src.astype(np.float64))


def test_old_style_header(tmp_path):
    """A type code outside 0..9 is read as a four-int header with float
    samples (src/FileTMP.cpp:210-214): here the first sample's bits (1.5f)
    sit where the type code would be."""
    im = (np.arange(24, dtype=np.float32) + 1.5).reshape(2, 1, 3, 4)
    p = str(tmp_path / "old.tmp")
    with open(p, "wb") as f:
        np.array([4, 3, 1, 2], np.int32).tofile(f)
        im.tofile(f)
    np.testing.assert_array_equal(G.tmp_read(p), im)


def test_read_errors(tmp_path):
    p = str(tmp_path / "short.tmp")
    with open(p, "wb") as f:
        np.array([4, 3, 1, 2, 0], np.int32).tofile(f)
        np.zeros(5, np.float32).tofile(f)
    with pytest.raises(G.GTError):
        G.tmp_read(p)
    with pytest.raises(G.GTError):
        G.tmp_read(str(tmp_path / "missing.tmp"))


def test_echo_server_socket_client():
    srv = G.NetServer(op=G.GT_NET_ECHO)
    try:
        rng = np.random.default_rng(1)
        for shape in [(3, 1, 17, 23), (1, 4, 5, 6), (2, 1, 1, 300)]:
            im = rng.normal(size=shape).astype(np.float32)
            with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as s:
                G.send_image(s, im)
                back = G.recv_image(s)
            assert np.array_equal(back, im)
        with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as s:
            s.sendall(np.array([0, 5, 1, 1], np.uint32).tobytes())     # zero width: dropped
            assert s.recv(16) == b""
        st = srv.stats()
        assert st["replied"] == 3 and st["failed"] == 1
    finally:
        srv.close()


def test_echo_server_byte_limits():
    """Unauthenticated headers: an image past max_bytes is dropped unread; one
    that would take the handlers past max_inflight bytes (input + output) is
    refused before anything is allocated, and accepted again once the bytes
    held drop."""
    srv = G.NetServer(op=G.GT_NET_ECHO, max_bytes=6000, max_inflight=10000)
    try:
        big = np.zeros((1, 1, 40, 40), np.float32)               # 6400 bytes > max_bytes
        with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as s:
            s.sendall(np.array([40, 40, 1, 1], np.uint32).tobytes())
            assert s.recv(16) == b""
        assert srv.stats()["failed"] == 1
        a = np.arange(1000, dtype=np.float32).reshape(1, 1, 25, 40)   # holds 2 x 4000 bytes
        b = np.ones((1, 1, 10, 50), np.float32)                       # needs 2 x 2000 more
        with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as sa:
            sa.sendall(np.array([40, 25, 1, 1], np.uint32).tobytes())   # header only: a handler waits, holding
            for _ in range(2000):
                if srv.stats()["inflight_bytes"] == 8000:
                    break
                time.sleep(0.005)
            assert srv.stats()["inflight_bytes"] == 8000
            with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as sb:
                assert dropped(sb, b)                           # refused
            assert srv.stats()["refused"] == 1
            sa.sendall(a.tobytes())
            assert np.array_equal(G.recv_image(sa), a)
        for _ in range(2000):
            if srv.stats()["inflight_bytes"] == 0:
                break
            time.sleep(0.005)
        with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as sb:
            G.send_image(sb, b)
            assert np.array_equal(G.recv_image(sb), b)
        st = srv.stats()
        assert st["replied"] == 2 and st["refused"] == 1 and st["failed"] == 1 and st["inflight_bytes"] == 0
    finally:
        srv.close()


def dropped(sock, im):
    """Sends im and reports whether the server closed the connection without
    a reply (the close may reach the client as EOF or as a reset)."""
    try:
        G.send_image(sock, im)
        return sock.recv(16) == b""
    except (ConnectionResetError, BrokenPipeError):
        return True


def test_echo_server_bind_address():
    """Loopback by default; an explicit address is honoured; a malformed one
    is refused."""
    srv = G.NetServer(op=G.GT_NET_ECHO, bind_addr="0.0.0.0")
    try:
        im = np.ones((1, 1, 2, 3), np.float32)
        with socket.create_connection(("127.0.0.1", srv.port), timeout=20) as s:
            G.send_image(s, im)
            assert np.array_equal(G.recv_image(s), im)
    finally:
        srv.close()
    with pytest.raises(G.GTError) as e:
        G.NetServer(op=G.GT_NET_ECHO, bind_addr="not-an-address")
    assert e.value.code == G.GT_EINVAL
    for kw, want in (({}, "0100007F"), ({"bind_addr": "0.0.0.0"}, "00000000")):
        srv = G.NetServer(op=G.GT_NET_ECHO, **kw)
        try:
            assert listen_address(srv.port) == want, kw
        finally:
            srv.close()


def listen_address(port):
    """Local address (hex, as /proc/net/tcp writes it) of the socket listening
    on `port`."""
    with open("/proc/net/tcp") as f:
        for line in f.readlines()[1:]:
            local, state = line.split()[1], line.split()[3]
            addr, p = local.split(":")
            if int(p, 16) == port and state == "0A":
                return addr
    return None


def free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def test_echo_server_reference_cli(tmp_path):
    """The unmodified reference CLI as the client: -send to the front end,
    -receive its reply on another port, -save: the file comes back intact."""
    if not os.path.exists(CLI):
        pytest.skip("reference CLI not built")
    reply = free_port()
    srv = G.NetServer(op=G.GT_NET_ECHO, reply_port=reply)
    try:
        im = np.random.default_rng(2).uniform(size=(3, 2, 30, 40)).astype(np.float32)
        src, dst = str(tmp_path / "a.tmp"), str(tmp_path / "b.tmp")
        G.tmp_write(src, im)
        r = subprocess.run([CLI, "-load", src, "-send", "127.0.0.1", str(srv.port), "-receive", str(reply),
                            "-save", dst], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        assert raw(dst) == raw(src)
        assert srv.stats()["replied"] == 1
    finally:
        srv.close()
