"""The C-ABI library loads on a CPU-only host and exports every entry point
include/*.h declares (no compute calls here: there is no GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wpic.h")
HEADER3D = os.path.join(ROOT, "include", "wpic3d.h")
HEADER_ADJ = os.path.join(ROOT, "include", "wpic_adj.h")
HEADER_IO = os.path.join(ROOT, "include", "wpic_io.h")
HEADER_APP = os.path.join(ROOT, "include", "wpic_app.h")
APP = os.path.join(ROOT, "wumingpic_amd", "app", "wpic_app2d")
LIB = os.path.join(ROOT, "wumingpic_amd", "libwpic.so")


def declared(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wpic_[a-z0-9_]+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(LIB), "build with `python -m wumingpic_amd.build` (or __graft_entry__.build())"


@pytest.mark.parametrize("header", [HEADER, HEADER3D, HEADER_ADJ, HEADER_IO, HEADER_APP])
def test_every_declared_symbol_is_exported(header):
    names = declared(header)
    assert len(names) >= (1 if header == HEADER_APP else 6)
    lib = C.CDLL(LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (wpic_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_python_binding_covers_the_header():
    from wumingpic_amd import _lib
    assert sorted(_lib.EXPORTS) == declared()
    assert sorted(_lib.EXPORTS3D) == declared(HEADER3D)
    assert sorted(_lib.EXPORTS_ADJ) == declared(HEADER_ADJ)
    assert sorted(_lib.EXPORTS_IO) == declared(HEADER_IO)
    assert sorted(_lib.EXPORTS_APP) == declared(HEADER_APP)


def test_missing_library_fails_loudly(monkeypatch):
    from wumingpic_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", os.path.join(ROOT, "no-such-lib.so"))
    with pytest.raises(ImportError):
        _lib.lib()


def test_invalid_parameters_rejected_without_gpu_work():
    """parameter validation happens before any device call"""
    from wumingpic_amd import _lib
    L = _lib.lib()
    p = _lib.Params2D()
    p.ndim, p.nsp = 5, 2
    h = C.c_void_p()
    assert L.wpic_create2d(C.byref(h), C.byref(p)) == _lib.E_INVAL
    assert b"ndim" in L.wpic_last_error()
    p.ndim = 6
    p.nxgs, p.nxge, p.nygs, p.nyge = 2, 65, 2, 65
    p.delx, p.delt, p.c, p.gfac = 1.0, 1.5, 1.0, 0.501
    p.capacity = 1000
    assert L.wpic_create2d(C.byref(h), C.byref(p)) == _lib.E_INVAL
    assert b"c*delt" in L.wpic_last_error()


def test_invalid_3d_parameters_rejected_without_gpu_work():
    from wumingpic_amd import _lib
    L = _lib.lib()
    p = _lib.Params3D()
    p.ndim, p.nsp = 6, 2
    h = C.c_void_p()
    assert L.wpic_create3d(C.byref(h), C.byref(p)) == _lib.E_INVAL
    assert b"ndim must be 7" in L.wpic_last_error()
    p.ndim = 7
    p.nxgs, p.nxge, p.nygs, p.nyge, p.nzgs, p.nzge = 2, 9, 2, 9, 2, 3
    p.delx, p.delt, p.c, p.gfac = 1.0, 0.5, 1.0, 0.501
    p.capacity = 1000
    assert L.wpic_create3d(C.byref(h), C.byref(p)) == _lib.E_INVAL
    assert b"4x3x3" in L.wpic_last_error()
    p.nzge = 9
    p.delx = 0.5
    assert L.wpic_create3d(C.byref(h), C.byref(p)) == _lib.E_INVAL
    assert b"delx" in L.wpic_last_error()


def test_app_driver_rejects_bad_configs_before_gpu_work(tmp_path):
    """load_config's errors (2d/proj/weibel/app.f90:160-250): a missing file, a file that is not
    JSON, a parameter section that names no single problem"""
    assert os.path.exists(APP)
    r = subprocess.run([APP, str(tmp_path / "none.json")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "does not exists" in r.stderr
    bad = tmp_path / "bad.json"
    bad.write_text("{ not json")
    r = subprocess.run([APP, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "failed to load" in r.stderr
    two = tmp_path / "two.json"
    two.write_text('{"config": {"max_it": 1}, "parameter": {"n_x": 8, "n_y": 8, "mass_ratio": 25.0, '
                   '"t_ani": 5.0, "nbg": 3, "ncs": 9}}')
    r = subprocess.run([APP, str(two)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "one problem" in r.stderr
