"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(gt_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "gt_gauss_transform" in names and "gt_last_error" in names
    assert len(names) >= 11


def test_library_exports_every_declared_symbol():
    from imagestack_amd import gtlattice
    if not os.path.exists(gtlattice.LIB_PATH):
        pytest.fail("libgtlattice.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(gtlattice.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from imagestack_amd import gtlattice
    assert set(gtlattice.EXPORTS) <= set(declared_functions())


def test_context_create_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from imagestack_amd import gtlattice
    with pytest.raises(gtlattice.GTError) as ei:
        gtlattice.Context(0)
    assert ei.value.code == 4
