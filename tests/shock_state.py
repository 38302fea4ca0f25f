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
