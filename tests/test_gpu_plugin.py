"""Drop-in through the reference's own plugin interface: the unmodified
ImageStack CLI (built from the reference sources into oracle/_ref) loads
imagestack_amd/libgt_plugin.so with -plugin and runs the GPU lattice behind
its own -gausstransform / -jointbilateral / -bilateral / -nlmeans / -nlmeans3d
operations."""
import os
import subprocess

import numpy as np
import pytest

from imagestack_amd import synth as S
from tests.util import FP32_REL_TOL, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "oracle", "_ref", "ImageStack")
PLUGIN = os.path.join(ROOT, "imagestack_amd", "libgt_plugin.so")


def need_cli():
    if not (os.path.exists(CLI) and os.path.exists(PLUGIN)):
        pytest.skip("reference CLI or plugin not built (reference sources absent at build time)")


def write_tmp(path, im):
    """ImageStack .tmp: int32 width, height, frames, channels, type 0, then
    planar float32 data by channel, frame, row (src/FileTMP.cpp)."""
    C, F, H, W = im.shape
    with open(path, "wb") as f:
        np.array([W, H, F, C, 0], np.int32).tofile(f)
        np.ascontiguousarray(im, np.float32).tofile(f)


def read_tmp(path):
    with open(path, "rb") as f:
        W, H, F, C, t = np.fromfile(f, np.int32, 5)
        assert t == 0
        return np.fromfile(f, np.float32, W * H * F * C).reshape(C, F, H, W)


def run(args, timeout=120):
    print("[cli]", " ".join(os.path.basename(a) if a.startswith("/") else a for a in args), flush=True)
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_plugin_runs_reference_style_tests():
    need_cli()
    out = run(["-plugin", PLUGIN, "-test", "gausstransform", "-test", "jointbilateral"])
    assert out.count("Passed") == 2, out
    assert "Failed" not in out


def test_plugin_bilateral_matches_reference_cli(tmp_path, oracle_lib):
    need_cli()
    im = S.rect_image(256, 192, 3, K=24, noise=0.05, seed=31)
    src = str(tmp_path / "in.tmp")
    write_tmp(src, im)
    cpu_path, gpu_path = str(tmp_path / "cpu.tmp"), str(tmp_path / "gpu.tmp")
    run(["-load", src, "-bilateral", "0.1", "8", "8", "0", "permutohedral", "-save", cpu_path])
    run(["-plugin", PLUGIN, "-load", src, "-bilateral", "0.1", "8", "8", "0", "permutohedral", "-save", gpu_path])
    cpu, gpu = read_tmp(cpu_path), read_tmp(gpu_path)
    o64 = oracle_lib.joint_bilateral(im, im, 8.0, 8.0, 0.0, 0.1, mode=oracle_lib.F64_LAT)
    assert rel_err(gpu, o64) < FP32_REL_TOL
    assert rel_err(gpu, cpu) < 2 * FP32_REL_TOL
    # AUTO picks the lattice for d = 5 (src/GaussTransform.cpp:553-564)
    auto_path = str(tmp_path / "auto.tmp")
    run(["-plugin", PLUGIN, "-load", src, "-bilateral", "0.1", "8", "-save", auto_path])
    assert rel_err(read_tmp(auto_path), o64) < FP32_REL_TOL


def test_plugin_gausstransform_recipe(tmp_path, oracle_lib):
    """The -gausstransform help recipe (src/GaussTransform.cpp:31-35): values
    [r,g,b,1], positions [x,y,r,g,b] pushed with -dup so slice and splat are
    distinct images (non-replay slice)."""
    need_cli()
    im = S.rect_image(96, 80, 3, K=10, noise=0.05, seed=32)
    W, H = 96, 80
    vals = S.homogeneous_values(im)
    xs = np.broadcast_to(np.arange(W, dtype=np.float32), (1, H, W))[None]
    ys = np.broadcast_to(np.arange(H, dtype=np.float32)[:, None], (1, H, W))[None]
    pos = np.concatenate([xs, ys, im]).astype(np.float32)
    vp, pp = str(tmp_path / "v.tmp"), str(tmp_path / "p.tmp")
    write_tmp(vp, vals)
    write_tmp(pp, pos)
    outs = {}
    for tag, plug in (("cpu", []), ("gpu", ["-plugin", PLUGIN])):
        o = str(tmp_path / f"{tag}.tmp")
        run(plug + ["-load", vp, "-load", pp, "-dup", "-gausstransform", "permutohedral", "4", "4", "0.1", "0.1",
                    "0.1", "-save", o])
        outs[tag] = read_tmp(o)
    sig = np.array([4, 4, 0.1, 0.1, 0.1], np.float32)
    inv = (np.float32(1) / sig).astype(np.float32)
    o64, _ = oracle_lib.gauss_transform(pos, vals, inv, slice_pos=pos, mode=oracle_lib.F64_LAT)
    assert rel_err(outs["gpu"].reshape(4, -1), o64) < FP32_REL_TOL
    assert rel_err(outs["gpu"], outs["cpu"]) < 2 * FP32_REL_TOL


def test_plugin_auto_grey_runs_the_grid(tmp_path, oracle_lib):
    """A grey -bilateral with AUTO has 3 position channels, so the reference
    picks GRID (src/GaussTransform.cpp:553-564); the plugin runs it on the
    GPU grid.  Checked against the FP64 grid oracle and the CPU CLI."""
    need_cli()
    im = S.rect_image(160, 120, 1, K=12, noise=0.05, seed=41)
    src = str(tmp_path / "grey.tmp")
    write_tmp(src, im)
    cpu_path, gpu_path, grid_path = (str(tmp_path / n) for n in ("cpu.tmp", "gpu.tmp", "grid.tmp"))
    run(["-load", src, "-bilateral", "0.1", "6", "-save", cpu_path])
    run(["-plugin", PLUGIN, "-load", src, "-bilateral", "0.1", "6", "-save", gpu_path])
    run(["-plugin", PLUGIN, "-load", src, "-bilateral", "0.1", "6", "6", "0", "grid", "-save", grid_path])
    o64 = oracle_lib.joint_bilateral_grid(im, im, 6.0, 6.0, 0.0, 0.1, mode=oracle_lib.F64_LAT)
    gpu = read_tmp(gpu_path)
    assert rel_err(gpu, o64) < FP32_REL_TOL
    assert rel_err(gpu, read_tmp(cpu_path)) < 2 * FP32_REL_TOL
    assert rel_err(read_tmp(grid_path), o64) < FP32_REL_TOL


def test_plugin_nlmeans(tmp_path):
    """-nlmeans / -nlmeans3d through the plugin: their reference-style tests
    (GPU result within 1e-4 of the reference's NLMeans::apply on the same
    rand() stream, and noise removed), then the CLI grammar on a .tmp image."""
    need_cli()
    out = run(["-plugin", PLUGIN, "-test", "nlmeans", "-test", "nlmeans3d"], timeout=300)
    assert out.count("Passed") == 2, out
    assert "Failed" not in out
    clean = S.rect_image(96, 72, 3, K=12, noise=0.0, seed=33)
    noisy = S.noisy_rect_image(96, 72, 3, K=12, amp=0.2, seed=33)
    src, dst = str(tmp_path / "noisy.tmp"), str(tmp_path / "out.tmp")
    write_tmp(src, noisy)
    run(["-plugin", PLUGIN, "-load", src, "-nlmeans", "1", "6", "16", "0.4", "-save", dst])
    res = read_tmp(dst)
    assert np.abs(res - clean).mean() < 0.5 * np.abs(noisy - clean).mean()


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def test_plugin_auto_zero_sigma_runs_on_the_gpu(tmp_path, oracle_lib, gpu_ctx):
    """AUTO with a zero spatial sigma: JointBilateral filters each row (or
    column) by itself and every piece makes the AUTO choice for its own shape
    (src/GaussTransform.cpp:516-538, :553-564).  A grey image's rows then
    take the grid, a 4-channel guide's rows the lattice; the plugin runs both
    on the GPU (its output is the library's bit for bit) and agrees with the
    reference CLI."""
    import torch
    need_cli()
    im = S.rect_image(200, 48, 1, K=12, noise=0.05, seed=51)
    src = str(tmp_path / "grey.tmp")
    write_tmp(src, im)
    cpu_path, gpu_path = str(tmp_path / "cpu.tmp"), str(tmp_path / "gpu.tmp")
    run(["-load", src, "-bilateral", "0.1", "6", "0", "-save", cpu_path])
    run(["-plugin", PLUGIN, "-load", src, "-bilateral", "0.1", "6", "0", "-save", gpu_path])
    gpu = read_tmp(gpu_path)
    lib_grid = gpu_ctx.joint_bilateral_grid(torch.from_numpy(im).cuda(), torch.from_numpy(im).cuda(), 6.0, 0.0,
                                            0.0, 0.1).cpu().numpy()
    assert same_bits(gpu, lib_grid)
    assert rel_err(gpu, read_tmp(cpu_path)) < 2 * FP32_REL_TOL
    guide = S.rect_image(200, 48, 4, K=12, noise=0.05, seed=52)
    image = S.uniform_image(200, 48, 3, seed=53)
    gp, ip = str(tmp_path / "guide.tmp"), str(tmp_path / "image.tmp")
    write_tmp(gp, guide)
    write_tmp(ip, image)
    run(["-load", gp, "-load", ip, "-jointbilateral", "0.15", "6", "0", "-save", cpu_path])
    run(["-plugin", PLUGIN, "-load", gp, "-load", ip, "-jointbilateral", "0.15", "6", "0", "-save", gpu_path])
    gpu = read_tmp(gpu_path)
    lib_lat = gpu_ctx.joint_bilateral(torch.from_numpy(image).cuda(), torch.from_numpy(guide).cuda(), 6.0, 0.0,
                                      0.0, 0.15).cpu().numpy()
    assert same_bits(gpu, lib_lat)
    o64 = oracle_lib.joint_bilateral(image, guide, 6.0, 0.0, 0.0, 0.15, mode=oracle_lib.F64_LAT)
    assert rel_err(gpu, o64) < FP32_REL_TOL
    assert rel_err(gpu, read_tmp(cpu_path)) < 2 * FP32_REL_TOL


def test_plugin_nonfinite_values_spread_as_in_reference(tmp_path):
    """NaN and Inf pixels: the plugin returns an image, with NaN/+Inf/-Inf
    exactly where the reference CLI puts them (src/Permutohedral.h:377)."""
    need_cli()
    im = S.rect_image(120, 90, 3, K=10, noise=0.05, seed=54)
    im[0, 0, 10, 20] = np.nan
    im[1, 0, 50, 70] = np.inf
    im[2, 0, 51, 70] = -np.inf
    src = str(tmp_path / "nan.tmp")
    write_tmp(src, im)
    cpu_path, gpu_path = str(tmp_path / "cpu.tmp"), str(tmp_path / "gpu.tmp")
    guide = S.rect_image(120, 90, 3, K=10, noise=0.05, seed=55)
    gp = str(tmp_path / "guide.tmp")
    write_tmp(gp, guide)
    run(["-load", gp, "-load", src, "-jointbilateral", "0.1", "5", "5", "0", "permutohedral", "-save", cpu_path])
    run(["-plugin", PLUGIN, "-load", gp, "-load", src, "-jointbilateral", "0.1", "5", "5", "0", "permutohedral",
         "-save", gpu_path])
    cpu, gpu = read_tmp(cpu_path), read_tmp(gpu_path)
    for f in (np.isnan, np.isposinf, np.isneginf):
        assert np.array_equal(f(gpu), f(cpu)), f.__name__
    assert np.isnan(gpu).any()
    fin = np.isfinite(cpu)
    assert rel_err(gpu[fin], cpu[fin]) < 2 * FP32_REL_TOL
    # a grey guide with AUTO takes the grid (3 position channels); the NaN/Inf
    # stay in the filtered image (a NaN position is an error here, and sends
    # the reference's grid index astray)
    grey = S.rect_image(120, 90, 1, K=10, noise=0.05, seed=57)
    gsrc = str(tmp_path / "grey.tmp")
    write_tmp(gsrc, grey)
    run(["-load", gsrc, "-load", src, "-jointbilateral", "0.1", "5", "-save", cpu_path])
    run(["-plugin", PLUGIN, "-load", gsrc, "-load", src, "-jointbilateral", "0.1", "5", "-save", gpu_path])
    cpu, gpu = read_tmp(cpu_path), read_tmp(gpu_path)
    for f in (np.isnan, np.isposinf, np.isneginf):
        assert np.array_equal(f(gpu), f(cpu)), f.__name__
    fin = np.isfinite(cpu)
    assert rel_err(gpu[fin], cpu[fin]) < 2 * FP32_REL_TOL


def test_plugin_lattice_beyond_16_dims_runs_the_reference(tmp_path):
    """-gausstransform permutohedral over 17 position channels (the reference
    allocates for any d, src/Permutohedral.h:229-240): beyond the library's
    16 the plugin hands the call to the reference's own code, so the output
    is the reference's bit for bit rather than an error."""
    need_cli()
    W, H = 24, 20
    s = S.Stream(56)
    pos = s.normal(17 * W * H).reshape(17, 1, H, W).astype(np.float32)
    vals = np.concatenate([s.uniform(2 * W * H).reshape(2, 1, H, W), np.ones((1, 1, H, W))]).astype(np.float32)
    vp, pp = str(tmp_path / "v.tmp"), str(tmp_path / "p.tmp")
    write_tmp(vp, vals)
    write_tmp(pp, pos)
    outs = {}
    for tag, plug in (("cpu", []), ("gpu", ["-plugin", PLUGIN])):
        o = str(tmp_path / f"{tag}.tmp")
        run(plug + ["-load", vp, "-load", pp, "-dup", "-gausstransform", "permutohedral", "1", "-save", o])
        outs[tag] = read_tmp(o)
    assert same_bits(outs["gpu"], outs["cpu"])
