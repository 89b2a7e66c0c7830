"""The C-ABI library: loads, exports every entry point include/porescale.h
declares, and reports errors through return codes (CPU only, no compute)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from porescale_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "porescale.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(psx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 28
    for s in syms:
        assert hasattr(L, s), f"{s} declared in porescale.h but not exported"
    assert set(syms) == set(_lib.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (psx_\w+)", out))
    assert set(syms) <= exported


def test_struct_layouts_match_header():
    # offsets the Fortran bind(C) types rely on
    assert C.sizeof(_lib.LBMParams) == 8 * 8 + 2 * 4 + 2 * 8 + 6 * 8 + 8 + 8 + 8 + 19 * 16 + 8 * 4
    assert _lib.LBMParams.precision.offset == _lib.LBMParams.Sb.offset + 19 * 8
    assert C.sizeof(_lib.LSMParams) == 7 * 8 + 8 * 4
    assert C.sizeof(_lib.LBMDiag) == (2 + 7 * 3 + 2 + 3 + 2 + 1) * 8 + 4 * 4 + 12 * 8


def _no_device():
    try:
        _lib.device_info()
        return False
    except _lib.PorescaleError:
        return True


@pytest.mark.skipif(not _no_device(), reason="a HIP device is present")
def test_errors_are_codes_not_aborts_without_gpu():
    from porescale_amd.io import LBMInput
    from porescale_amd.lbm import LBM
    with pytest.raises(_lib.PorescaleError) as ei:
        LBM(LBMInput(ncx=4, ncy=4, ncz=4), np.ones((4, 4, 4), np.int32))
    assert ei.value.code == _lib.PSX_EHIP
    assert "device" in str(ei.value).lower()


def test_argument_validation_precedes_device_use():
    L = _lib.lib()
    p = _lib.LBMParams()
    p.dt, p.rr0, p.rb0, p.idir = 1.0, 1.0, 1.0, 7
    h = C.c_void_p()
    ind = np.ones(64, np.int32)
    rc = L.psx_lbm_create(C.byref(p), 4, 4, 4, _lib.ptr(ind), C.byref(h))
    assert rc == _lib.PSX_EINVAL and b"idir" in L.psx_last_error()
    q = _lib.LSMParams()
    q.rs, q.dt, q.precision = 1.0, 0.01, 9
    rc = L.psx_lsm_create(C.byref(q), 4, 4, 4, _lib.ptr(ind), None, None, None, None, None, C.byref(h))
    assert rc == _lib.PSX_EINVAL and b"precision" in L.psx_last_error()


def test_fortran_drivers_built_and_report_missing_input(tmp_path):
    import subprocess
    fort = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "porescale_amd", "fortran")
    for name in ("psx_lbm_drv", "psx_lsm_drv"):
        exe = os.path.join(fort, name)
        assert os.path.exists(exe), f"{name} not built"
        r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "INPUT_PARAMETERS" in r.stdout


def test_noalloc_limits_checked_before_device_use():
    """psx_lsm_stress_noalloc refuses volumes beyond its fixed scratch and
    bad load cases before any device call (the pointers are not touched)."""
    L = _lib.lib()
    buf = np.zeros(16)
    p = _lib.ptr(buf)
    rc = L.psx_lsm_stress_noalloc(300, 300, 300, p, p, p, 0, 0.01, p, None)
    assert rc == _lib.PSX_EINVAL and b"scratch" in L.psx_last_error()
    assert L.psx_lsm_stress_noalloc(4, 4, 4, p, p, p, 6, 0.01, p, None) == _lib.PSX_EINVAL
    assert L.psx_lsm_stress_noalloc(4, 4, 4, None, p, p, 0, 0.01, p, None) == _lib.PSX_EINVAL
