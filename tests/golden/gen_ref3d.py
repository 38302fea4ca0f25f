"""Write golden vectors from the reference's OWN 3D particle__solv,
particle__solv_vay, sort__bucket and mom_calc (3d/common/particle.f90:52-419,
3d/common/sort.f90:40-88, 3d/common/mom_calc.f90:49-333),
compiled from the reference sources by `make -C oracle ref` and run at one
OpenMP thread.

    python tests/golden/gen_ref3d.py        (needs oracle/_ref/ref_golden3d)

Outputs tests/golden/ref3d_push.npz, ref3d_vay.npz, ref3d_sort.npz and ref3d_mom.npz: the
inputs, the parameters and the reference's outputs.  The fixtures are data
only; tests/test_oracle3d.py checks the oracle against them.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import pyoracle as po  # noqa: E402
import states3d as S  # noqa: E402


def params_dict(p):
    return dict(ndim=p.ndim, np=p.np, nsp=p.nsp, nxgs=p.nxgs, nxge=p.nxge, nygs=p.nygs, nyge=p.nyge,
                nzgs=p.nzgs, nzge=p.nzge, nys=p.nys, nye=p.nye, nzs=p.nzs, nze=p.nze, delx=p.delx,
                delt=p.delt, c=p.c, q=np.array([p.q[k] for k in range(p.nsp)]),
                r=np.array([p.r[k] for k in range(p.nsp)]))


def push_case(tmp):
    rng = np.random.default_rng(20261203)
    nx, ny, nz, n_ppc = 8, 6, 5, 4
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=64, delt=0.5, q=q, r=r)
    counts = [rng.integers(1, 7, size=(nx, ny, nz)) for _ in range(2)]
    up, cumcnt, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=(0.05, 0.4), drift=(0.0, 0.5),
                                           colocate=False, counts=counts)
    # edge positions (a cell's lower face, just below the upper face) and a
    # few strongly relativistic momenta
    up[0:3, 0, 0, 0, 0] = [float(p.nxgs), float(p.nys) + 0.999999999999, float(p.nzs)]
    up[2, 1, 2, 3, 1] = np.nextafter(float(p.nzs + 4), 0.0)
    up[3:6, 2, 3, 1, 1] = [40.0, -25.0, 7.5]
    up[3:6, 0, 4, 2, 0] = [-3.0, 2.0, -1.0]
    # inactive slots hold arbitrary bits: the id component is copied over the
    # whole capacity (particle.f90:227-231)
    nmax = int(np2.max())
    up[6, nmax:] = rng.standard_normal(up[6, nmax:].shape)
    uf = S.random_fields3d(p, rng, scale=0.3, b0=0.2)
    gp = po.run_ref3d(os.path.join(tmp, "push"), 1, p, up, uf=uf, cumcnt=cumcnt)
    np.savez_compressed(os.path.join(HERE, "ref3d_push.npz"), up=up, uf=uf, cumcnt=cumcnt, np2=np2,
                        gp=gp, **params_dict(p))


def vay_case(tmp):
    """particle__solv_vay with strong fields (|fac1 B| ~ 1 for the electrons)"""
    rng = np.random.default_rng(20261207)
    nx, ny, nz, n_ppc = 8, 6, 5, 4
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=64, delt=0.5, q=q, r=r)
    counts = [rng.integers(1, 7, size=(nx, ny, nz)) for _ in range(2)]
    up, cumcnt, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=(0.05, 0.4), drift=(0.0, 0.5),
                                           colocate=False, counts=counts)
    up[3:6, 2, 3, 1, 1] = [40.0, -25.0, 7.5]
    nmax = int(np2.max())
    up[6, nmax:] = rng.standard_normal(up[6, nmax:].shape)
    bscale = 1.0 / abs(q[1] * 0.25)
    uf = S.random_fields3d(p, rng, scale=0.3 * bscale, b0=0.5 * bscale)
    gp = po.run_ref3d(os.path.join(tmp, "vay"), 4, p, up, uf=uf, cumcnt=cumcnt)
    np.savez_compressed(os.path.join(HERE, "ref3d_vay.npz"), up=up, uf=uf, cumcnt=cumcnt, np2=np2,
                        gp=gp, **params_dict(p))


def sort_case(tmp):
    rng = np.random.default_rng(8)
    nx, ny, nz = 8, 6, 5
    p = po.make_params3d(nx, ny, nz, np_cap=64, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    up = np.zeros((p.ndim, p.np, ny, nz, 2), order="F")
    np2 = rng.integers(20, 64, size=(ny, nz, 2)).astype(np.int32).copy(order="F")
    for s in range(2):
        for k in range(nz):
            for j in range(ny):
                n = int(np2[j, k, s])
                x = p.nxgs + nx * rng.random(n)
                x[:3] = np.floor(x[:3])
                x[3:6] = np.nextafter(np.floor(x[3:6]) + 1.0, 0.0)
                up[0, :n, j, k, s] = x
                up[1, :n, j, k, s] = p.nys + j + rng.random(n)
                up[2, :n, j, k, s] = p.nzs + k + rng.random(n)
                up[3:6, :n, j, k, s] = rng.standard_normal((3, n))
                up[6, :n, j, k, s] = S.ids_as_f64(1 + (j + 100 * k) * 100 + s * 10 ** 6 + np.arange(n))
    gp, cc = po.run_ref3d(os.path.join(tmp, "sort"), 2, p, up, np2=np2)
    np.savez_compressed(os.path.join(HERE, "ref3d_sort.npz"), up=up, np2=np2, gp=gp, cumcnt=cc,
                        **params_dict(p))


def mom_case(tmp):
    """mom_calc__accl then mom_calc__nvt on a sorted 3D state."""
    rng = np.random.default_rng(20261205)
    nx, ny, nz, n_ppc = 8, 6, 5, 4
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=64, delt=0.5, q=q, r=r)
    counts = [rng.integers(1, 7, size=(nx, ny, nz)) for _ in range(2)]
    up, cumcnt, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=(0.05, 0.4), drift=(0.0, 0.5),
                                           colocate=False, counts=counts)
    up[0:3, 0, 0, 0, 0] = [float(p.nxgs), float(p.nys) + 0.999999999999, float(p.nzs) + 0.5]
    up[3:6, 0, 4, 2, 0] = [-3.0, 2.0, -1.0]
    up[3:6, 2, 3, 1, 1] = [40.0, -25.0, 7.5]
    uf = S.random_fields3d(p, rng, scale=0.3, b0=0.2)
    gp, mom = po.run_ref3d(os.path.join(tmp, "mom"), 3, p, up, np2=np2, uf=uf, cumcnt=cumcnt)
    np.savez_compressed(os.path.join(HERE, "ref3d_mom.npz"), up=up, uf=uf, cumcnt=cumcnt, np2=np2,
                        gp=gp, mom=mom, **params_dict(p))


if __name__ == "__main__":
    if not po.have_ref3d():
        sys.exit("oracle/_ref/ref_golden3d is missing: run `make -C oracle ref` first")
    with tempfile.TemporaryDirectory() as tmp:
        push_case(tmp)
        vay_case(tmp)
        sort_case(tmp)
        mom_case(tmp)
    print("wrote", *(os.path.join(HERE, f"ref3d_{k}.npz") for k in ("push", "vay", "sort", "mom")))
