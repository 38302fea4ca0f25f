"""The reference 3D app's call sequence (3d/proj/weibel/app.f90:103-108 and the
moment output sequence :122-124), written in Fortran against the 3D shim
modules (particle, field, sort, mom_calc, boundary_periodic with
__particle_yz), run on the GPU: in parity mode the state after the loop and
the moments equal the 3D oracle's bit for bit, with host mirroring after every
call and in resident mode."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
import states3d as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "wumingpic_amd", "fortran", "shim_step3d_driver")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _state(seed=5):
    rng = np.random.default_rng(seed)
    n_ppc, nx, ny, nz = 4, 10, 6, 5
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=(0.02, 0.2), drift=(0.0, 0.5), colocate=True)
    uf = S.periodic_fill_fields3d(p, S.random_fields3d(p, rng, scale=0.01))
    return p, up, uf, cc, np2


@pytest.mark.parametrize("resident", [0, 1])
def test_fortran_3d_app_loop_on_shim_matches_oracle(resident):
    assert os.path.exists(DRIVER), "build with __graft_entry__.build()"
    p, up, uf, cc, np2 = _state()
    nsteps = 4
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "params.txt"), "w") as f:
            f.write(f"{nsteps} 1 {resident} {p.ndim} {p.np} {p.nsp} {p.nxgs} {p.nxge} {p.nygs} {p.nyge} "
                    f"{p.nzgs} {p.nzge}\n")
            f.write(f"{p.delx!r} {p.delt!r} {p.c!r} {p.gfac!r}\n")
            f.write(" ".join(repr(float(p.q[k])) for k in range(p.nsp)) + "\n")
            f.write(" ".join(repr(float(p.r[k])) for k in range(p.nsp)) + "\n")
        up.ravel(order="F").tofile(os.path.join(tmp, "up.bin"))
        uf.ravel(order="F").tofile(os.path.join(tmp, "uf.bin"))
        cc.ravel(order="F").tofile(os.path.join(tmp, "cumcnt.bin"))
        np2.ravel(order="F").tofile(os.path.join(tmp, "np2.bin"))
        print(f"3D shim driver: {p.nx}x{p.ny}x{p.nz}, {nsteps} steps, resident={resident}", flush=True)
        r = subprocess.run([DRIVER, tmp], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        u_dev = np.fromfile(os.path.join(tmp, "up.out")).reshape(up.shape, order="F")
        f_dev = np.fromfile(os.path.join(tmp, "uf.out")).reshape(uf.shape, order="F")
        c_dev = np.fromfile(os.path.join(tmp, "cumcnt.out"), dtype=np.int32).reshape(cc.shape, order="F")
        n_dev = np.fromfile(os.path.join(tmp, "np2.out"), dtype=np.int32).reshape(np2.shape, order="F")
        m_dev = np.fromfile(os.path.join(tmp, "mom.out")).reshape((7, p.nx + 2, p.ny + 2, p.nz + 2, p.nsp),
                                                                  order="F")
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, nsteps)
    assert np.array_equal(n_dev, np2)
    assert np.array_equal(c_dev, cc)
    for s in range(p.nsp):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(u_dev[:, :n, j, k, s]), _bits(up[:, :n, j, k, s])), (s, j, k)
    assert np.array_equal(_bits(f_dev), _bits(uf))
    # the moment output sequence after the loop (3d/proj/weibel/app.f90:122-124)
    o.mom_accl(gp, up, uf, cc)
    mom = o.mom_nvt(gp, np2)
    o.bc_mom(mom)
    assert np.array_equal(_bits(m_dev), _bits(mom))
