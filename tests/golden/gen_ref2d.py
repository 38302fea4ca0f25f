"""Write golden vectors from the reference's OWN 2D particle__solv,
particle__solv_vay, sort__bucket and mom_calc (2d/common/particle.f90:48-315,
2d/common/sort.f90:36-82, 2d/common/mom_calc.f90:48-249),
compiled from the reference sources by `make -C oracle ref` and run at one
OpenMP thread.

    python tests/golden/gen_ref2d.py        (needs oracle/_ref/ref_golden2d)

Outputs tests/golden/ref2d_push.npz, ref2d_vay.npz, ref2d_sort.npz and ref2d_mom.npz: the
inputs, the parameters and the reference's outputs.  The fixtures are data
only; tests/test_oracle_golden.py checks the oracle against them.
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
import states2d as S  # noqa: E402


def params_dict(p):
    return dict(ndim=p.ndim, np=p.np, nsp=p.nsp, nxgs=p.nxgs, nxge=p.nxge, nygs=p.nygs,
                nyge=p.nyge, nys=p.nys, nye=p.nye, delx=p.delx, delt=p.delt, c=p.c,
                q=np.array([p.q[k] for k in range(p.nsp)]), r=np.array([p.r[k] for k in range(p.nsp)]))


def push_case(tmp):
    rng = np.random.default_rng(20261202)
    nx, ny, n_ppc = 12, 10, 8
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=160, delt=0.5, q=q, r=r)
    counts = [rng.integers(4, 13, size=(nx, ny)) for _ in range(2)]
    up, cumcnt, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.05, 0.4), drift=(0.0, 0.5),
                                         colocate=False, counts=counts)
    # edge positions: exactly on the left face of a cell (dh = -1/2), just
    # below the right face, and a few strongly relativistic momenta
    up[0, 0, 0, 0] = float(p.nxgs)
    up[1, 0, 0, 0] = float(p.nys) + 0.999999999999
    up[0, 3, 2, 1] = np.nextafter(float(p.nxgs + 1), 0.0)
    up[2:5, 5, 3, 1] = [40.0, -25.0, 7.5]
    up[2:5, 6, 4, 0] = [-3.0, 2.0, -1.0]
    # inactive slots hold arbitrary bits: the reference copies the id slot
    # over the whole capacity (particle.f90:173-177)
    up[5, np2.max():, :, :] = rng.standard_normal(up[5, np2.max():, :, :].shape)
    uf = S.random_fields(p, rng, scale=0.3, b0=0.2)
    gp = po.run_ref2d(os.path.join(tmp, "push"), 1, p, up, uf=uf, cumcnt=cumcnt)
    np.savez_compressed(os.path.join(HERE, "ref2d_push.npz"), up=up, uf=uf, cumcnt=cumcnt, np2=np2,
                        gp=gp, **params_dict(p))


def vay_case(tmp):
    """particle__solv_vay on the push case's kind of state, with strong fields so that the implicit
    rotation matters (|tau| up to about 1), and a particle at the E x B drift velocity"""
    rng = np.random.default_rng(20261206)
    nx, ny, n_ppc = 12, 10, 8
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=160, delt=0.5, q=q, r=r)
    counts = [rng.integers(4, 13, size=(nx, ny)) for _ in range(2)]
    up, cumcnt, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.05, 0.4), drift=(0.0, 0.5),
                                         colocate=False, counts=counts)
    up[0, 0, 0, 0] = float(p.nxgs)
    up[0, 3, 2, 1] = np.nextafter(float(p.nxgs + 1), 0.0)
    up[2:5, 5, 3, 1] = [40.0, -25.0, 7.5]
    up[2:5, 6, 4, 0] = [-3.0, 2.0, -1.0]
    up[5, np2.max():, :, :] = rng.standard_normal(up[5, np2.max():, :, :].shape)
    # |fac1 B| ~ 1 for the electrons: fac1 = q/r dt/2
    bscale = 1.0 / abs(q[1] * 0.25)
    uf = S.random_fields(p, rng, scale=0.3 * bscale, b0=0.5 * bscale)
    gp = po.run_ref2d(os.path.join(tmp, "vay"), 4, p, up, uf=uf, cumcnt=cumcnt)
    np.savez_compressed(os.path.join(HERE, "ref2d_vay.npz"), up=up, uf=uf, cumcnt=cumcnt, np2=np2,
                        gp=gp, **params_dict(p))


def sort_case(tmp):
    rng = np.random.default_rng(7)
    nx, ny = 12, 10
    p = po.make_params(nx, ny, np_cap=160, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    up = np.zeros((p.ndim, p.np, p.ny, p.nsp), order="F")
    np2 = rng.integers(60, 150, size=(ny, 2)).astype(np.int32).copy(order="F")
    for s in range(2):
        for j in range(ny):
            n = int(np2[j, s])
            x = p.nxgs + nx * rng.random(n)
            # integer faces and values just below them
            x[:5] = np.floor(x[:5])
            x[5:10] = np.nextafter(np.floor(x[5:10]) + 1.0, 0.0)
            up[0, :n, j, s] = x
            up[1, :n, j, s] = p.nys + j + rng.random(n)
            up[2:5, :n, j, s] = rng.standard_normal((3, n))
            up[5, :n, j, s] = S.ids_as_f64(1 + j * 1000 + s * 100000 + np.arange(n))
    gp, cc = po.run_ref2d(os.path.join(tmp, "sort"), 2, p, up, np2=np2)
    np.savez_compressed(os.path.join(HERE, "ref2d_sort.npz"), up=up, np2=np2, gp=gp, cumcnt=cc,
                        **params_dict(p))


def mom_case(tmp):
    """mom_calc__accl then mom_calc__nvt on a sorted state (2d/proj/weibel/app.f90:121-122)."""
    rng = np.random.default_rng(20261204)
    nx, ny, n_ppc = 12, 10, 8
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=160, delt=0.5, q=q, r=r)
    counts = [rng.integers(4, 13, size=(nx, ny)) for _ in range(2)]
    up, cumcnt, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.05, 0.4), drift=(0.0, 0.5),
                                         colocate=False, counts=counts)
    # a particle on a cell's left face, one on a cell centre (dx = 0 in the
    # CIC weights), one just below the domain's top face, a relativistic one
    up[0, 0, 0, 0] = float(p.nxgs)
    up[0:2, 1, 2, 0] = [float(p.nxgs) + 0.5 + float(int(up[0, 1, 2, 0] - p.nxgs)), float(p.nys + 2) + 0.5]
    up[1, 2, ny - 1, 1] = np.nextafter(float(p.nye + 1), 0.0)
    up[2:5, 3, 4, 1] = [40.0, -25.0, 7.5]
    uf = S.random_fields(p, rng, scale=0.3, b0=0.2)
    gp, mom = po.run_ref2d(os.path.join(tmp, "mom"), 3, p, up, np2=np2, uf=uf, cumcnt=cumcnt)
    np.savez_compressed(os.path.join(HERE, "ref2d_mom.npz"), up=up, uf=uf, cumcnt=cumcnt, np2=np2,
                        gp=gp, mom=mom, **params_dict(p))


if __name__ == "__main__":
    if not po.have_ref():
        sys.exit("oracle/_ref/ref_golden2d is missing: run `make -C oracle ref` first")
    with tempfile.TemporaryDirectory() as tmp:
        push_case(tmp)
        vay_case(tmp)
        sort_case(tmp)
        mom_case(tmp)
    print("wrote", *(os.path.join(HERE, f"ref2d_{k}.npz") for k in ("push", "vay", "sort", "mom")))
