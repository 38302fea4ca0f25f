"""The Fortran shim archives export the reference's module procedures under
the reference's module names: 2d/common (particle, field, sort, mom_calc,
boundary_periodic, plus the 2D projects' boundary modules) and 3d/common
(particle, field, sort, mom_calc, boundary_periodic with __particle_yz).
Symbol names follow flang's module mangling, _QM<module>P<procedure>."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORT = os.path.join(ROOT, "wumingpic_amd", "fortran")

COMMON = {
    "particle": ["particle__init", "particle__solv", "particle__solv_vay"],
    "field": ["field__init", "field__fdtd_i"],
    "sort": ["sort__init", "sort__bucket"],
    "mom_calc": ["mom_calc__init", "mom_calc__accl", "mom_calc__nvt"],
}
BC2D = ["__init", "__particle_x", "__particle_y", "__dfield", "__curre", "__phi", "__mom"]
BC3D = ["__init", "__particle_x", "__particle_yz", "__dfield", "__curre", "__phi", "__mom"]


def _symbols(archive):
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm is not available")
    out = subprocess.run([nm, archive], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def _expect(mods):
    return {f"_QM{m}P{p}" for m, procs in mods.items() for p in procs}


def test_2d_shim_archive_exports_2d_common_procedures():
    archive = os.path.join(FORT, "libwpic_fortran.a")
    assert os.path.exists(archive), "build with __graft_entry__.build()"
    mods = dict(COMMON)
    mods["boundary_periodic"] = ["boundary_periodic" + s for s in BC2D]
    mods["boundary_reconnection"] = ["boundary_reconnection" + s for s in BC2D]
    mods["boundary_shock"] = ["boundary_shock" + s for s in BC2D] + ["boundary_shock__injection"]
    missing = _expect(mods) - _symbols(archive)
    assert not missing, sorted(missing)


def test_3d_shim_archive_exports_3d_common_procedures():
    archive = os.path.join(FORT, "libwpic_fortran3d.a")
    assert os.path.exists(archive), "build with __graft_entry__.build()"
    mods = dict(COMMON)
    mods["boundary_periodic"] = ["boundary_periodic" + s for s in BC3D]
    syms = _symbols(archive)
    missing = _expect(mods) - syms
    assert not missing, sorted(missing)
    # the 3D archive carries no 2D-only boundary modules
    assert not any(s.startswith("_QMboundary_reconnection") for s in syms)
    assert os.path.exists(os.path.join(FORT, "shim_step3d_driver"))
