"""Checks that pin the oracle where no reference build can: closed forms,
conservation laws and exact solves.  Each is chosen so that a plausible slip in
the restated algorithm (a dropped term, a sign, a shifted index, a transposed
operand) fails it.

  round-down wrap      exact rational arithmetic (fractions)
  Boris push           gyration angle 2 atan(q dt B / 2 m gamma), |u| kept;
                       pure E kick of 2 fac1 E; x += u dt / gamma
  Esirkepov deposit    discrete continuity rho1 - rho0 + dt div J = 0 per cell
                       for Jx, Jy; Jz = q v_z s0 s0 for a particle at rest in x-y
  CG solve             FFT solve of the periodic operator; dense solve with walls
  whole step           Gauss's law residual stays at round-off; particle count,
                       ids and row membership conserved; sortedness
"""
from fractions import Fraction
import math

import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S


def _rd_exact(fr):
    d = float(fr)
    if Fraction(d) > fr:
        d = math.nextafter(d, -math.inf)
    return d


def test_round_down_arithmetic_exact():
    L = po.lib()
    rng = np.random.default_rng(0)
    cases = [(2.0 - 2.0 ** -52, 1024.0), (2.0 - 2.0 ** -52, 64.0), (1.9999999999999998, 64.0),
             (1026.0 - 2.0 ** -42, -1024.0), (0.1, 0.2), (1.0, -1.0)]
    cases += [(float(a), float(b)) for a, b in zip(rng.uniform(1, 3, 200), rng.choice([64.0, -64.0, 1024.0], 200))]
    for a, b in cases:
        assert L.or_add_rd(a, b) == _rd_exact(Fraction(a) + Fraction(b)), (a, b)
    # the wrap edge case: just below nxgs=2 wraps to just below nxge+1
    assert L.or_add_rd(2.0 - 2.0 ** -52, 1024.0) == 1025.9999999999998
    for a, b in zip(rng.uniform(0.5, 3, 100), rng.uniform(0.5, 3, 100)):
        assert L.or_mul_rd(a, b) == _rd_exact(Fraction(a) * Fraction(b))
        assert L.or_div_rd(a, b) == _rd_exact(Fraction(a) / Fraction(b))


def _uniform_state(B=(0, 0, 0), E=(0, 0, 0), n_ppc=6, seed=3, u_scale=0.3):
    rng = np.random.default_rng(seed)
    p = po.make_params(8, 8, np_cap=3 * n_ppc * 8, delt=0.5, q=(0.3, -0.7), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(u_scale, u_scale), drift=(0.0, 0.0), colocate=False)
    uf = np.zeros((6, p.nx + 4, p.ny + 4), order="F")
    for k in range(3):
        uf[k] = B[k]
        uf[3 + k] = E[k]
    return p, up, uf, cc, np2


def _active(a, np2):
    return np.concatenate([a[:, :np2[j, s], j, s] for s in range(a.shape[3]) for j in range(a.shape[2])], axis=1)


def test_boris_uniform_b_gyration():
    B0 = 0.8
    p, up, uf, cc, np2 = _uniform_state(B=(0, 0, B0))
    gp = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv(gp, up, uf, cc)
    for s in range(2):
        u0 = np.concatenate([up[2:5, :np2[j, s], j, s] for j in range(p.ny)], axis=1)
        u1 = np.concatenate([gp[2:5, :np2[j, s], j, s] for j in range(p.ny)], axis=1)
        gam = np.sqrt(1 + (u0 ** 2).sum(0))
        fac1 = p.q[s] / p.r[s] * 0.5 * p.delt
        # |u| conserved, u_z untouched
        np.testing.assert_allclose(np.linalg.norm(u1, axis=0), np.linalg.norm(u0, axis=0), rtol=2e-15)
        np.testing.assert_array_equal(u1[2], u0[2])
        ang = np.arctan2(u0[0] * u1[1] - u0[1] * u1[0], u0[0] * u1[0] + u0[1] * u1[1])
        np.testing.assert_allclose(ang, -2 * np.arctan(fac1 * B0 / gam), rtol=1e-12, atol=1e-15)


def test_boris_uniform_e_kick_and_move():
    E0 = np.array([0.05, -0.02, 0.03])
    p, up, uf, cc, np2 = _uniform_state(E=tuple(E0))
    gp = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv(gp, up, uf, cc)
    for s in range(2):
        u0 = np.concatenate([up[:5, :np2[j, s], j, s] for j in range(p.ny)], axis=1)
        u1 = np.concatenate([gp[:5, :np2[j, s], j, s] for j in range(p.ny)], axis=1)
        fac1 = p.q[s] / p.r[s] * 0.5 * p.delt
        np.testing.assert_allclose(u1[2:5] - u0[2:5], np.broadcast_to((2 * fac1 * E0)[:, None], u1[2:5].shape),
                                   rtol=1e-11, atol=1e-16)
        gam = np.sqrt(1 + (u1[2:5] ** 2).sum(0))
        np.testing.assert_allclose(u1[0], u0[0] + u1[2] * p.delt / gam, rtol=0, atol=4e-16 * 16)
        np.testing.assert_allclose(u1[1], u0[1] + u1[3] * p.delt / gam, rtol=0, atol=4e-16 * 16)


def _rho_centres(p, xy, q):
    """charge at cell centres with the quadratic spline, periodic fold"""
    nx, ny = p.nx, p.ny
    rho = np.zeros((nx, ny))
    x, y = xy
    i0 = np.floor(x).astype(int)
    j0 = np.floor(y).astype(int)
    dx = x - 0.5 - i0
    dy = y - 0.5 - j0
    sx = [0.5 * (0.5 - dx) ** 2, 0.75 - dx ** 2, 0.5 * (0.5 + dx) ** 2]
    sy = [0.5 * (0.5 - dy) ** 2, 0.75 - dy ** 2, 0.5 * (0.5 + dy) ** 2]
    for a in range(3):
        for b in range(3):
            np.add.at(rho, ((i0 - 1 + a - p.nxgs) % nx, (j0 - 1 + b - p.nys) % ny), q * sx[a] * sy[b])
    return rho


def test_deposit_discrete_continuity():
    rng = np.random.default_rng(4)
    p = po.make_params(12, 10, np_cap=90, delt=0.5, q=(0.37, -0.21), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles(p, 5, rng, u_scale=(0.1, 0.1), drift=(0.0, 0.0), colocate=False)
    gp = up.copy(order="F")
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            gp[0, :n, j, s] += rng.uniform(-0.49, 0.49, n)
            gp[1, :n, j, s] += rng.uniform(-0.49, 0.49, n)
    uj = np.zeros((3, p.nx + 4, p.ny + 4), order="F")
    o = po.Oracle2D(p)
    o.ele_cur(uj, up, gp, cc)
    o.bc_curre(uj)
    drho = np.zeros((p.nx, p.ny))
    for s in range(2):
        xy0 = np.concatenate([up[:2, :np2[j, s], j, s] for j in range(p.ny)], axis=1)
        xy1 = np.concatenate([gp[:2, :np2[j, s], j, s] for j in range(p.ny)], axis=1)
        drho += _rho_centres(p, xy1, p.q[s]) - _rho_centres(p, xy0, p.q[s])
    # interior of uj in array coordinates: 2..nx+1; face i+1 is the next column
    jx, jy = uj[0], uj[1]
    divj = (jx[3:p.nx + 3, 2:p.ny + 2] - jx[2:p.nx + 2, 2:p.ny + 2]) + \
           (jy[2:p.nx + 2, 3:p.ny + 3] - jy[2:p.nx + 2, 2:p.ny + 2])
    resid = drho + p.delt * divj
    assert np.abs(drho).max() > 1e-3
    assert np.abs(resid).max() < 1e-15 * 64


def test_deposit_jz_of_particle_at_rest_in_plane():
    rng = np.random.default_rng(8)
    p = po.make_params(8, 8, np_cap=40, delt=0.5, q=(0.5, -0.5), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles(p, 1, rng, u_scale=(0.0, 0.0), drift=(0.0, 0.0), colocate=False,
                                     counts=[np.zeros((8, 8), int), np.zeros((8, 8), int)])
    # one electron at rest in x-y moving in z
    up[:, 0, 3, 1] = [p.nxgs + 4.3, p.nys + 3.6, 0.0, 0.0, 0.7, 0.0]
    np2[3, 1] = 1
    cc[:, 3, 1] = 0
    cc[5:, 3, 1] = 1
    gp = up.copy(order="F")
    uj = np.zeros((3, p.nx + 4, p.ny + 4), order="F")
    po.Oracle2D(p).ele_cur(uj, up, gp, cc)
    assert np.abs(uj[:2]).max() == 0.0
    vz = 0.7 / np.sqrt(1 + 0.49)
    dx, dy = 4.3 - 4 - 0.5, 3.6 - 3 - 0.5
    sx = np.array([0.5 * (0.5 - dx) ** 2, 0.75 - dx ** 2, 0.5 * (0.5 + dx) ** 2])
    sy = np.array([0.5 * (0.5 - dy) ** 2, 0.75 - dy ** 2, 0.5 * (0.5 + dy) ** 2])
    expect = -0.5 * vz * np.outer(sx, sy)
    # cell (i, j) = (nxgs+4, nys+3) sits at array column 4+2, row 3+2
    np.testing.assert_allclose(uj[2, 5:8, 4:7], expect, rtol=1e-14)
    assert abs(uj[2].sum() - (-0.5 * vz)) < 1e-15


def test_cg_matches_fft_solve_periodic():
    rng = np.random.default_rng(12)
    p = po.make_params(16, 12, np_cap=10, delt=0.5, q=(1.0, -1.0), r=(25.0, 1.0))
    o = po.Oracle2D(p)
    f = np.zeros(5)
    po.lib().or2d_field_consts(p, f)
    f4, f5 = f[3], f[4]
    o.gkl[:] = rng.standard_normal(o.gkl.shape)
    o.df[:] = 0.0
    iters = o.cgm()
    assert min(iters) >= 1
    kx = 2 * np.pi * np.fft.fftfreq(p.nx)
    ky = 2 * np.pi * np.fft.fftfreq(p.ny)
    sym = f4 - 2 * np.cos(kx)[:, None] - 2 * np.cos(ky)[None, :]
    for l in range(3):
        b = f5 * o.gkl[l]
        phi = np.real(np.fft.ifft2(np.fft.fft2(b) / sym))
        got = o.df[l, 2:p.nx + 2, 2:p.ny + 2]
        assert np.abs(got - phi).max() <= 1e-5 * np.abs(phi).max()


def test_cg_matches_dense_solve_with_walls():
    rng = np.random.default_rng(13)
    p = po.make_params(9, 6, np_cap=10, delt=0.5, q=(1.0, -1.0), r=(25.0, 1.0), bc=po.BC_WALLX)
    o = po.Oracle2D(p)
    f = np.zeros(5)
    po.lib().or2d_field_consts(p, f)
    f4, f5 = f[3], f[4]
    o.gkl[:] = rng.standard_normal(o.gkl.shape)
    o.df[:] = 0.0
    o.cgm()
    nx, ny = p.nx, p.ny
    for l in range(3):
        # ghost values as linear maps of the interior, taken from the oracle's
        # own boundary routine applied to unit vectors
        A = np.zeros((nx * ny, nx * ny))
        for k in range(nx * ny):
            phi = np.zeros((nx + 2, ny + 2), order="F")
            phi[1 + k % nx, 1 + k // nx] = 1.0
            o.bc_phi(phi, l + 1)
            lap = (f4 * phi[1:-1, 1:-1] - phi[:-2, 1:-1] - phi[2:, 1:-1] - phi[1:-1, :-2] - phi[1:-1, 2:])
            A[:, k] = lap.ravel(order="F")
        b = (f5 * o.gkl[l]).ravel(order="F")
        phi = np.linalg.solve(A, b).reshape((nx, ny), order="F")
        got = o.df[l, 2:nx + 2, 2:ny + 2]
        assert np.abs(got - phi).max() <= 1e-5 * np.abs(phi).max()


def test_wall_phi_boundary_table():
    """the reconnection ghost table as written in boundary_reconnection.f90:557-577"""
    p = po.make_params(7, 5, np_cap=10, bc=po.BC_WALLX)
    o = po.Oracle2D(p)
    base = np.arange(1.0, (p.nx + 2) * (p.ny + 2) + 1).reshape((p.nx + 2, p.ny + 2), order="F")
    for l, expect in ((1, lambda a, j: (-a[1, j], -a[p.nx - 2, j])), (2, lambda a, j: (a[2, j], a[p.nx - 1, j]))):
        a = base.copy(order="F")
        o.bc_phi(a, l)
        for j in range(1, p.ny + 1):
            lo, hi = expect(a, j)
            assert a[0, j] == lo and a[p.nx + 1, j] == hi


def _weibel_oracle_state(nx, ny, n_ppc, seed=20261202):
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=True)
    uf = np.zeros((6, nx + 4, ny + 4), order="F")
    return p, up, uf, cc, np2


def test_full_step_conservation_and_gauss_law():
    p, up, uf, cc, np2 = _weibel_oracle_state(16, 16, 8)
    o = po.Oracle2D(p)
    ids0 = np.sort(_active(up, np2)[5].view(np.int64))
    G0 = o.gauss(uf, up, np2)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 15)
    G1 = o.gauss(uf, up, np2)
    scale = 4 * np.pi * abs(p.q[1]) * 8
    assert np.abs(G0).max() < 1e-13 * scale
    assert np.abs(G1).max() < 1e-12 * scale
    assert np.abs(uf).max() > 1e-8  # the beams did drive currents and fields
    assert np.array_equal(np.sort(_active(up, np2)[5].view(np.int64)), ids0)
    assert np2.sum() == 2 * 8 * 16 * 16
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            x, y = up[0, :n, j, s], up[1, :n, j, s]
            assert (np.floor(y).astype(int) == p.nys + j).all()
            cells = np.floor(x).astype(int)
            assert (np.diff(cells) >= 0).all()
            assert cc[-1, j, s] == n
            assert np.array_equal(np.bincount(cells - p.nxgs, minlength=p.nx), np.diff(cc[:, j, s]))


def test_step_thread_count_independent_at_the_cell_level():
    """the OpenMP baseline build gives the same cells as one thread after a
    few steps (only reduction order differs)"""
    p, up, uf, cc, np2 = _weibel_oracle_state(16, 16, 8)
    st = [a.copy(order="F") for a in (up, uf, cc, np2)]
    o1 = po.Oracle2D(p, threads=1)
    o1.step(st[0], np.zeros_like(up, order="F"), st[1], st[2], st[3], 4)
    o4 = po.Oracle2D(p, threads=4)
    o4.step(up, np.zeros_like(up, order="F"), uf, cc, np2, 4)
    po.lib().or_set_threads(1)
    assert np.array_equal(cc, st[2])
    assert np.abs(uf - st[1]).max() <= 1e-12 * np.abs(st[1]).max()
