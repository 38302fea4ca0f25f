"""An inflow state for the shock boundaries (2d/proj/shock/app.f90:395-470 with the box fully
expanded): particles in cells nxgs+1..nxge-2 with thermal momenta plus the inflow u0 along x,
the upstream field B = b0 (cos theta, sin theta cos phi, sin theta sin phi) and E = -v0 x B / c."""
import numpy as np

from oracle import pyoracle as po
import states2d as S


def shock_state(nx, ny, n_ppc, seed, u0=-0.3, theta=np.pi / 3, phi=np.pi / 2, vth=(0.02, 0.1), delt=0.5):
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=delt, q=q, r=r, bc=po.BC_SHOCK, u0=u0)
    counts = np.zeros((nx, ny), dtype=np.int64)
    counts[1:nx - 2, :] = rng.integers(n_ppc // 2, 3 * n_ppc // 2, size=(nx - 3, ny))
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=vth, drift=(0.0, 0.0), colocate=True,
                                     counts=[counts, counts])
    for s in range(2):
        for j in range(ny):
            up[2, :np2[j, s], j, s] += u0
    b0 = 0.5 / abs(q[0])
    v0 = u0 / np.sqrt(1 + u0 * u0)
    uf = np.zeros((6, nx + 4, ny + 4), order="F")
    uf[0] = b0 * np.cos(theta)
    uf[1] = b0 * np.sin(theta) * np.cos(phi)
    uf[2] = b0 * np.sin(theta) * np.sin(phi)
    uf[4] = +v0 * uf[2]
    uf[5] = -v0 * uf[1]
    uf += 0.01 * b0 * rng.standard_normal(uf.shape)
    return p, np.asfortranarray(up), np.asfortranarray(uf), cc, np2


def wider_box(p, uf, cc, extra, rng):
    """The same state in a box `extra` cells wider whose active edge nxe is the narrow box's
    nxge (the shock's expanding box, 2d/proj/shock/app.f90:276-277): the narrow planes fill the
    first columns, the columns beyond get unrelated values (a field there must not reach the
    active box), the cumcnt entries beyond nxe+1 carry each row's total."""
    pb = po.make_params(p.nx + extra, p.ny, np_cap=p.np, delt=p.delt, q=[p.q[k] for k in range(p.nsp)],
                        r=[p.r[k] for k in range(p.nsp)], bc=p.bc, u0=p.u0)
    pb.nxe = p.nxge
    ufb = np.asfortranarray(rng.standard_normal((6, pb.nx + 4, p.ny + 4)) * np.abs(uf).max())
    ufb[:, :p.nx + 4, :] = uf
    ccb = np.zeros((pb.nx + 1, p.ny, p.nsp), dtype=np.int32, order="F")
    ccb[:p.nx + 1] = cc
    ccb[p.nx + 1:] = cc[-1][None]
    return pb, ufb, ccb
