"""The C-ABI library loads and exports every symbol include/hm_bem.h declares
(no compute: runs without a GPU)."""
import os
import re

import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "hm_bem.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("hm_fingreen3d_batch", "hm_infgreen3d_batch", "hm_sglintbd_batch", "hm_assemble",
                 "hm_factor", "hm_solve", "hm_forces", "hm_zgetrf", "hm_zgetrs"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import ctypes

    import hm_bem
    lib = hm_bem.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing
    assert sorted(hm_bem.EXPORTS) == _declared()
    assert isinstance(lib, ctypes.CDLL)


def test_dispersion_entry_runs_on_host(oracle):
    # the only entry point that needs no device
    import hm_bem
    import numpy as np
    np.testing.assert_array_equal(hm_bem.dispersion(0.8, 50.0, 9.807), oracle.dispersion(0.8, 50.0))


def test_fortran_layer_built():
    """the drop-in module object and its Fortran test caller, which links the
    reference-named wrappers against libhm_bem.so (run on the GPU by
    tests/test_gpu_fortran.py); every wrapper the caller uses resolves"""
    import subprocess
    assert os.path.exists(os.path.join(ROOT, "hams-mrel_amd", "lib", "hm_bem_f.o"))
    exe = os.path.join(ROOT, "tests", "fortran", "hm_fortran_check")
    assert os.path.exists(exe)
    r = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert "libhm_bem.so" in r.stdout and "not found" not in r.stdout, r.stdout
