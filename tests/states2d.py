"""Seeded 2D states in the reference's own array layout, for tests.

These are numpy builders for inputs; they are not the product's initial-
condition generator (that runs on the device).  Layouts follow
2d/proj/weibel/app.f90:278-283: up(ndim,np,nys:nye,nsp), uf(6,...), cumcnt, np2.
"""
import numpy as np


def ids_as_f64(pid):
    """Particle ids as the reference stores them: int64 -pid bit-cast into a
    real(8) slot (2d/proj/weibel/app.f90:468)."""
    return (-np.asarray(pid, dtype=np.int64)).view(np.float64)


def weibel_consts(n_ppc, omega_pe=0.1, mass_ratio=25.0, delx=1.0, c=1.0):
    """q, r as 2d/proj/weibel/app.f90:292-303 computes them (sigma_e = 0)."""
    pi = np.pi
    wpe = omega_pe
    wpi = wpe / np.sqrt(mass_ratio)
    r = np.array([mass_ratio, 1.0])
    q = np.array([+np.sqrt(r[0] / (4 * pi * n_ppc / delx ** 2)) * wpi,
                  -np.sqrt(r[1] / (4 * pi * n_ppc / delx ** 2)) * wpe])
    return q, r


def sorted_particles(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=True,
                     counts=None):
    """Particles uniformly random inside each cell, rows sorted by cell, with
    cumcnt / np2 consistent.  Electrons (isp=2) get +/- drift along x,
    alternating per slot.  Returns up, cumcnt, np2 (reference layout)."""
    nx, ny, nsp = p.nx, p.ny, p.nsp
    up = np.zeros((p.ndim, p.np, ny, nsp), order="F")
    cumcnt = np.zeros((nx + 1, ny, nsp), dtype=np.int32, order="F")
    np2 = np.zeros((ny, nsp), dtype=np.int32, order="F")
    pid0 = 0
    pos = None
    for s in range(nsp):
        if counts is None:
            cnt = np.full((nx, ny), n_ppc, dtype=np.int64)
        else:
            cnt = counts[s]
        if not (colocate and s > 0 and pos is not None):
            pos = []
            for j in range(ny):
                cj = cnt[:, j]
                tot = int(cj.sum())
                ic = np.repeat(np.arange(nx), cj)
                x = (p.nxgs + ic + rng.random(tot)) * p.delx
                y = (p.nys + j + rng.random(tot)) * p.delx
                pos.append((x, y, cj))
        for j in range(ny):
            x, y, cj = pos[j]
            tot = len(x)
            up[0, :tot, j, s] = x
            up[1, :tot, j, s] = y
            u = rng.standard_normal((3, tot)) * u_scale[min(s, len(u_scale) - 1)]
            d = drift[min(s, len(drift) - 1)]
            if d != 0.0:
                sign = np.where(np.arange(tot) % 2 == 0, 1.0, -1.0)
                u[0] += sign * d
            up[2:5, :tot, j, s] = u
            if p.ndim == 6:
                up[5, :tot, j, s] = ids_as_f64(pid0 + 1 + np.arange(tot))
            pid0 += tot
            np2[j, s] = tot
            cumcnt[0, j, s] = 0
            cumcnt[1:, j, s] = np.cumsum(cj)
    return up, cumcnt, np2


def random_fields(p, rng, scale=0.2, b0=0.0):
    uf = (rng.standard_normal((6, p.nx + 4, p.ny + 4)) * scale).copy(order="F")
    uf[2] += b0
    return uf


def periodic_fill_fields(p, uf):
    """Make ghost layers of uf consistent with periodicity (for physical runs)."""
    nx, ny = p.nx, p.ny
    # interior index range in array coords: 2..nx+1, 2..ny+1
    uf[:, :, 0:2] = uf[:, :, ny:ny + 2]
    uf[:, :, ny + 2:ny + 4] = uf[:, :, 2:4]
    uf[:, 0:2, :] = uf[:, nx:nx + 2, :]
    uf[:, nx + 2:nx + 4, :] = uf[:, 2:4, :]
    return uf
