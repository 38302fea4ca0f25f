"""The 3D Weibel app on the GPU (wumingpic_amd/app/wpic_app3d over include/wpic_app.h): its lattice
initial state against its formulas (3d/proj/weibel/app.f90 set_initial_condition and
set_particle_ids), and a whole driver run in parity mode against the oracle run from the same
initial state: the particle and moment files byte for byte, energy.dat with one kinetic column
per species, the final restart."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "wumingpic_amd", "app", "wpic_app3d")
PI = 3.141592653589793


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _consts(n_ppc, mass_ratio=25.0, sigma_e=0.1, omega_pe=0.1):
    """3d/proj/weibel/app.f90:296-308"""
    wpe = omega_pe
    wge = omega_pe * np.sqrt(sigma_e)
    wpi = wpe / np.sqrt(mass_ratio)
    wgi = wge / mass_ratio
    r = (mass_ratio, 1.0)
    q = (+np.sqrt(r[0] / (4.0 * PI * n_ppc)) * wpi, -np.sqrt(r[1] / (4.0 * PI * n_ppc)) * wpe)
    return q, r, r[0] * 1.0 / q[0] * wgi, (wpe, wpi, wge, wgi)


def _pic(nx, ny, nz, q, r, mode, cap, delt):
    from wumingpic_amd.pic3d import Pic3D
    return Pic3D(nx, ny, nz, list(q), list(r), delt=delt, mode=mode, capacity=cap)


def test_weibel3d_initial_state():
    from wumingpic_amd import MODE_PARITY
    nx, ny, nz, n0 = 12, 4, 3, 5
    q, r, b0, _ = _consts(n0)
    print("weibel3d initial state", flush=True)
    with _pic(nx, ny, nz, q, r, MODE_PARITY, n0 * nx * ny * nz, 1.0) as sim:
        sim.ic_weibel(3, n0, 0.02, 0.1, 5.0, b0)
        up, np2, cc = sim.download_up(n0 * nx)
        uf = sim.download_fields()
    assert np.all(np2 == n0 * nx)
    ii = np.arange(1, n0 * nx + 1)
    x = (2 + nx * (ii - 5e-1) / (n0 * nx)) * 1.0
    for s in range(2):
        for k in range(nz):
            for j in range(ny):
                assert np.array_equal(up[0, :, j, k, s], x)
                assert np.all(np.floor(up[1, :, j, k, s]) == 2 + j) and np.all(np.floor(up[2, :, j, k, s]) == 2 + k)
                col = j + ny * k
                assert np.array_equal(up[6, :, j, k, s].view(np.int64), -(col * n0 * nx + ii))
    assert np.array_equal(up[1:3, :, :, :, 0], up[1:3, :, :, :, 1])
    assert np.all(uf[2] == b0) and np.all(uf[[0, 1, 3, 4, 5]] == 0)
    u = up[5, :, :, :, 1].ravel()
    assert abs(u.std() / (5.0 * 0.1) - 1) < 0.1


def test_driver3d_run_matches_the_oracle():
    from wumingpic_amd import MODE_PARITY
    from wumingpic_amd.paraio import ParaIO
    nx, ny, nz, n0, steps = 16, 6, 4, 4, 4
    par = {"num_process": 1, "n_ppc": n0, "n_x": nx, "n_y": ny, "n_z": nz, "mass_ratio": 25.0, "sigma_e": 0.1,
           "omega_pe": 0.1, "v_the": 0.1, "v_thi": 0.02, "t_ani": 5.0, "seed": 7}
    cfg = {"config": {"verbose": 0, "datadir": ".", "max_elapsed": 1.0e6, "max_it": steps, "intvl_ptcl": steps,
                      "intvl_mom": 2, "intvl_orb": 100}, "parameter": par}
    q, r, b0, (wpe, wpi, wge, wgi) = _consts(n0)
    np_cap, delt = 3 * n0 * nx, 1.0
    print("driver3d run", flush=True)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "config.json")
        with open(path, "w") as f:
            json.dump(cfg, f)
        rr = subprocess.run([APP, path, "parity"], capture_output=True, text=True, timeout=120, cwd=d)
        assert rr.returncode == 0, rr.stdout + rr.stderr
        files = {n: open(os.path.join(d, n), "rb").read() for n in os.listdir(d) if n.endswith((".json", ".raw"))}
        edat = open(os.path.join(d, "energy.dat")).read().splitlines()
    assert "%07d_restart.json" % (steps + 1) in files
    with _pic(nx, ny, nz, q, r, MODE_PARITY, n0 * nx * ny * nz, delt) as sim:
        sim.ic_weibel(7, n0, 0.02, 0.1, 5.0, b0)
        up, np2, cc = sim.download_up(np_cap)
        uf = sim.download_fields()
    p = po.make_params3d(nx, ny, nz, np_cap=np_cap, delt=delt, q=q, r=r)
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, steps)
    with tempfile.TemporaryDirectory() as t:
        io = ParaIO(ndim=7, np_cap=np_cap, nsp=2, nxgs=2, nxge=nx + 1, nygs=2, nyge=ny + 1, nzgs=2, nzge=nz + 1,
                    delx=1.0, delt=delt, c=1.0, q=list(q), r=list(r), dir=t + "/")
        io.ptcl(up, uf, np2, steps)
        io.param(n0, wpe, wpi, wge, wgi, 0.02, 0.1, "init_param")
        o.mom_accl(gp, up, uf, cc)
        mom = o.mom_nvt(gp, np2)
        o.bc_mom(mom)
        io.mom(mom, uf, steps)
        io.close()
        for name in ("%07d_ptcl" % steps, "%07d_mom" % steps, "init_param"):
            for ext in (".json", ".raw"):
                assert files[name + ext] == open(os.path.join(t, name + ext), "rb").read(), name + ext
    # '(f10.2, 5(1x, e12.5))': time, ion and electron kinetic, E, B, total
    assert len(edat) == 3
    line = edat[-1]
    assert len(line) == 10 + 5 * 13
    vals = [float(v) for v in line.split()]
    assert vals[0] == steps * delt
    e = o.energy(up, uf, np2)
    assert np.isclose(vals[1] + vals[2], e[0], rtol=2e-4)
    assert np.allclose(vals[3:5], e[1:], rtol=1e-4) and np.isclose(vals[5], e.sum(), rtol=1e-4)
    assert vals[1] > 0 and vals[2] > 0
