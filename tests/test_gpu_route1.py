"""Route 1: the in-process GaussTransform::PERMUTOHEDRAL_GPU method case.

The build patches the reference's GaussTransform.{h,cpp} with a sed recipe in
imagestack_amd/Makefile (src/GaussTransform.h:8, src/GaussTransform.cpp:134
and :566) and links
imagestack_amd/host/route1_check.cpp with the reference's objects and
libgtlattice.so.  That program compares GaussTransform::apply and
JointBilateral::apply (with and without a zero spatial sigma) between the
reference's PERMUTOHEDRAL and the patched PERMUTOHEDRAL_GPU on the same images.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "imagestack_amd", "_route1", "route1_check")


@pytest.mark.gpu
def test_route1_method_case_matches_reference():
    if not os.path.exists(BIN):
        pytest.skip("route-1 check not built (built only where the reference sources are present)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    errs = [float(words[i + 1]) for i, w in enumerate(words) if w in ("gausstransform", "jointbilateral",
                                                                       "jointbilateral_rows")]
    assert len(errs) == 3 and max(errs) < 1e-5, r.stdout
