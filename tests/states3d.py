"""Seeded 3D states in the reference's own array layout, for tests.

Layouts follow 3d/proj/weibel/app.f90:282-286: up(7,np,nys:nye,nzs:nze,nsp)
with the id bit-cast into component 7, uf(6,nxgs-2:nxge+2,nys-2:nye+2,
nzs-2:nze+2), cumcnt(nxgs:nxge+1,nys:nye,nzs:nze,nsp), np2(nys:nye,nzs:nze,nsp).
"""
import numpy as np

from states2d import ids_as_f64, weibel_consts  # noqa: F401  (re-exported)


def sorted_particles3d(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=True, counts=None):
    """Particles uniformly random inside each cell, every (j, k) column sorted
    by x cell, cumcnt / np2 consistent.  Electrons (last species) get +/- drift
    along x alternating per slot.  Returns up, cumcnt, np2."""
    nx, ny, nz, nsp = p.nx, p.ny, p.nz, p.nsp
    up = np.zeros((p.ndim, p.np, ny, nz, nsp), order="F")
    cumcnt = np.zeros((nx + 1, ny, nz, nsp), dtype=np.int32, order="F")
    np2 = np.zeros((ny, nz, nsp), dtype=np.int32, order="F")
    pid0 = 0
    pos = None
    for s in range(nsp):
        cnt = np.full((nx, ny, nz), n_ppc, dtype=np.int64) if counts is None else counts[s]
        if not (colocate and s > 0 and pos is not None):
            pos = {}
            for k in range(nz):
                for j in range(ny):
                    cj = cnt[:, j, k]
                    tot = int(cj.sum())
                    ic = np.repeat(np.arange(nx), cj)
                    x = (p.nxgs + ic + rng.random(tot)) * p.delx
                    y = (p.nys + j + rng.random(tot)) * p.delx
                    z = (p.nzs + k + rng.random(tot)) * p.delx
                    pos[j, k] = (x, y, z, cj)
        for k in range(nz):
            for j in range(ny):
                x, y, z, cj = pos[j, k]
                tot = len(x)
                up[0, :tot, j, k, s] = x
                up[1, :tot, j, k, s] = y
                up[2, :tot, j, k, s] = z
                u = rng.standard_normal((3, tot)) * u_scale[min(s, len(u_scale) - 1)]
                d = drift[min(s, len(drift) - 1)]
                if d != 0.0:
                    u[0] += np.where(np.arange(tot) % 2 == 0, 1.0, -1.0) * d
                up[3:6, :tot, j, k, s] = u
                up[6, :tot, j, k, s] = ids_as_f64(pid0 + 1 + np.arange(tot))
                pid0 += tot
                np2[j, k, s] = tot
                cumcnt[1:, j, k, s] = np.cumsum(cj)
    return up, cumcnt, np2


def random_fields3d(p, rng, scale=0.2, b0=0.0):
    uf = (rng.standard_normal((6, p.nx + 4, p.ny + 4, p.nz + 4)) * scale).copy(order="F")
    uf[2] += b0
    return uf


def periodic_fill_fields3d(p, uf):
    """Make every ghost layer of uf consistent with periodicity in x, y and z."""
    nx, ny, nz = p.nx, p.ny, p.nz
    uf[:, :, :, 0:2] = uf[:, :, :, nz:nz + 2]
    uf[:, :, :, nz + 2:nz + 4] = uf[:, :, :, 2:4]
    uf[:, :, 0:2, :] = uf[:, :, ny:ny + 2, :]
    uf[:, :, ny + 2:ny + 4, :] = uf[:, :, 2:4, :]
    uf[:, 0:2, :, :] = uf[:, nx:nx + 2, :, :]
    uf[:, nx + 2:nx + 4, :, :] = uf[:, 2:4, :, :]
    return uf
