"""Pins the oracle's implicit field update (field__fdtd_i: the RHS gkl, the
three solves, the E update, the ghost fills and uf += df;
2d/common/field.f90:124-184, 3d/common/field.f90:127-206) by what the
mathematics fixes, not by restating its loops.

1. One Fourier mode through one step.  The update is meant to be Maxwell's
   equations on the staggered grid, time-centred with weight g (gfac):
       B' = B - c dt curl(g E' + (1 - g) E)
       E' = E + c dt curl(g B' + (1 - g) B) - 4 pi dt J
   Every difference in it is centred between two staggered nodes, so on a
   plane wave exp(i k.x) each d/dx acts as i K_x with K_x = 2 sin(k_x dx/2)/dx.
   The test solves that 6x6 linear system for one mode with numpy and checks
   the oracle's real-space step against it, fields and ghosts, with and
   without a prescribed current.  The staggering is the reference's layout
   (particle.f90:74-80 centring, field.f90's index offsets): in units of dx,
       2D  Bx (1/2,0)   By (0,1/2)   Bz (0,0)     Ex (0,1/2)   Ey (1/2,0)   Ez (1/2,1/2)
       3D  Bx (1/2,0,0) By (0,1/2,0) Bz (0,0,1/2) Ex (0,1/2,1/2) Ey (1/2,0,1/2) Ez (1/2,1/2,0)
   and each current component sits on its E component.  A flipped sign or a
   wrong factor (f1, f2, f3, gfac, 4 pi dt) anywhere in gkl or in df(4:6), a
   wrong neighbour index, or a missing ghost fill changes the step by far more
   than the tolerance; the last test checks that the mode model itself would
   see such mistakes.
2. div B stays at round-off over a particle run (the RHS is a curl plus the
   Laplacian of B, and the Laplacian commutes with div), and Gauss's law holds
   while div E itself is far from zero.

The CG runs with a relative tolerance of 1e-13 here (the reference's 1e-6 is
a stopping rule, not part of the scheme).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S
import states3d as S3
from fourier_mode import OFF2, OFF3, cross_matrix, current_of, fields_of, mode_amplitudes, mode_step

@pytest.fixture(autouse=True)
def tight_cg():
    old = po.set_cg_err(1e-13)
    yield
    po.set_cg_err(old)


def _mode2d(m, seed):
    nx, ny = 16, 12
    p = po.make_params(nx, ny, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    k = np.array([2 * np.pi * m[0] / nx, 2 * np.pi * m[1] / ny, 0.0])
    K = 2 * np.sin(k / 2)
    B, E, J = mode_amplitudes(K, np.random.default_rng(seed))  # Kz = 0 leaves Bz free
    return p, k, K, B, E, J


def _grid2d(p):
    # node index relative to nxgs / nys on the stored box nxgs-2..nxge+2
    return (p.nx + 4, p.ny + 4), (-2, -2)


@pytest.mark.parametrize("m", [(1, 2), (3, -1), (0, 1), (2, 0)])
@pytest.mark.parametrize("with_current", [False, True])
def test_fourier_mode_steps_2d(m, with_current):
    p, k, K, B, E, J = _mode2d(m, seed=abs(7 + m[0] * 10 + m[1]))
    if not with_current:
        J = np.zeros(3, dtype=complex)
    shape, lo = _grid2d(p)
    offs = OFF2
    uf = fields_of(B, E, k[:2], offs, shape, lo)
    o = po.Oracle2D(p)
    scale = np.abs(uf).max()
    for step in range(3):  # later steps start the CG from the previous dB
        o.uj[...] = current_of(J, k[:2], offs, shape, lo)
        o.field_from_current(uf)
        B, E = mode_step(K, B, E, J, p.c, p.delt, p.gfac)
        want = fields_of(B, E, k[:2], offs, shape, lo)
        err = np.abs(uf - want).max()
        assert err <= 1e-11 * scale, (step, err)


def _mode3d(m, seed):
    nx, ny, nz = 8, 6, 4
    p = po.make_params3d(nx, ny, nz, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    k = np.array([2 * np.pi * m[0] / nx, 2 * np.pi * m[1] / ny, 2 * np.pi * m[2] / nz])
    K = 2 * np.sin(k / 2)
    B, E, J = mode_amplitudes(K, np.random.default_rng(seed))
    return p, k, K, B, E, J


@pytest.mark.parametrize("m", [(1, 2, 1), (2, -1, 0), (0, 0, 1), (3, 2, -1)])
@pytest.mark.parametrize("with_current", [False, True])
def test_fourier_mode_steps_3d(m, with_current):
    p, k, K, B, E, J = _mode3d(m, seed=abs(3 + m[0] + 5 * m[1] + 25 * m[2]))
    if not with_current:
        J = np.zeros(3, dtype=complex)
    shape, lo = (p.nx + 4, p.ny + 4, p.nz + 4), (-2, -2, -2)
    uf = fields_of(B, E, k, OFF3, shape, lo)
    o = po.Oracle3D(p)
    scale = np.abs(uf).max()
    for step in range(2):
        o.uj[...] = current_of(J, k, OFF3, shape, lo)
        o.field_from_current(uf)
        B, E = mode_step(K, B, E, J, p.c, p.delt, p.gfac)
        want = fields_of(B, E, k, OFF3, shape, lo)
        err = np.abs(uf - want).max()
        assert err <= 1e-11 * scale, (step, err)


def test_mode_model_sees_plausible_mistakes():
    """the mode check above is not blind: a flipped sign of either curl or of
    J, gfac replaced by 1/2, a missing factor 4 pi, or a neighbour shifted by
    one node (wrong staggering of one component) each moves the one-step
    result of the 2D test mode by more than 1e6 times its tolerance"""
    p, k, K, B, E, J = _mode2d((1, 2), seed=11)
    c, dt, g = p.c, p.delt, p.gfac
    B1, E1 = mode_step(K, B, E, J, c, dt, g)
    ref = np.concatenate([B1, E1])
    scale = np.abs(np.concatenate([B, E])).max()

    def far(b, e):
        return np.abs(np.concatenate([b, e]) - ref).max() > 1e6 * 1e-11 * scale

    C = cross_matrix(1j * K.astype(complex))
    I3 = np.eye(3)

    def variant(sb, se, sj, gg, fpi):
        A = np.block([[I3, sb * c * dt * gg * C], [-se * c * dt * gg * C, I3]])
        rhs = np.concatenate([B - sb * c * dt * (1 - gg) * (C @ E),
                              E + se * c * dt * (1 - gg) * (C @ B) - sj * fpi * dt * J])
        x = np.linalg.solve(A, rhs)
        return x[:3], x[3:]

    assert far(*variant(-1, 1, 1, g, 4 * np.pi))
    assert far(*variant(1, -1, 1, g, 4 * np.pi))
    assert far(*variant(1, 1, -1, g, 4 * np.pi))
    assert far(*variant(1, 1, 1, 0.5, 4 * np.pi))
    assert far(*variant(1, 1, 1, g, 1.0))
    # Ez read one node off in y: its y difference symbol picks up exp(i ky)
    Kbad = K.copy().astype(complex)
    Kbad[1] = K[1] * np.exp(1j * k[1])
    Cb = cross_matrix(1j * Kbad)
    A = np.block([[I3, c * dt * g * Cb], [-c * dt * g * C, I3]])
    rhs = np.concatenate([B - c * dt * (1 - g) * (Cb @ E), E + c * dt * (1 - g) * (C @ B) - 4 * np.pi * dt * J])
    x = np.linalg.solve(A, rhs)
    assert far(x[:3], x[3:])


def _div_b2(uf):
    # div B at the Bz nodes (i, j): Bx(i+1/2,j) - Bx(i-1/2,j) + By(i,j+1/2) - By(i,j-1/2)
    bx, by = uf[0], uf[1]
    return (bx[2:-2, 2:-2] - bx[1:-3, 2:-2]) + (by[2:-2, 2:-2] - by[2:-2, 1:-3])


def _div_e2(uf):
    # div E at the cell centres, where the charge of the shape function lives
    ex, ey = uf[3], uf[4]
    return (ex[3:-1, 2:-2] - ex[2:-2, 2:-2]) + (ey[2:-2, 3:-1] - ey[2:-2, 2:-2])


def test_div_b_and_gauss_over_a_particle_run_2d():
    """20 steps of the Weibel beams: div B stays at round-off (a sign or factor
    slip in any one curl or current term of gkl breaks it), and Gauss's law
    div E = 4 pi rho holds where div E itself is large"""
    nx, ny, n_ppc = 16, 16, 8
    rng = np.random.default_rng(5)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    # ions and electrons not co-located: rho and div E are O(1) from the start
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=False)
    uf = np.zeros((6, nx + 4, ny + 4), order="F")
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    G0 = o.gauss(uf, up, np2)
    assert np.abs(G0).max() > 1e-3  # E = 0 against a lumpy charge: G = -4 pi rho
    # start from a field that satisfies Gauss: one solve of the particles'
    # own dynamics does not provide it, so only the change of G is checked
    dive0 = _div_e2(uf)
    o.step(up, gp, uf, cc, np2, 20)
    G1 = o.gauss(uf, up, np2)
    dive1 = _div_e2(uf)
    divb = _div_b2(uf)
    bmax = np.abs(uf[0:3]).max()
    assert bmax > 1e-6
    assert np.abs(divb).max() <= 1e-12 * bmax
    # the residual G keeps its initial value: dG/dt = 0 to round-off
    assert np.abs(G1 - G0).max() <= 1e-11 * np.abs(G0).max()
    # and div E is far from zero and far from its start: the check is not vacuous
    assert np.abs(dive1 - dive0).max() > 1e3 * np.abs(G1 - G0).max()


def test_div_b_over_a_particle_run_3d():
    nx, ny, nz, n_ppc = 8, 8, 6, 4
    rng = np.random.default_rng(9)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S3.sorted_particles3d(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=False)
    uf = np.zeros((6, nx + 4, ny + 4, nz + 4), order="F")
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    G0 = o.gauss(uf, up, np2)
    o.step(up, gp, uf, cc, np2, 8)
    G1 = o.gauss(uf, up, np2)
    bx, by, bz = uf[0], uf[1], uf[2]
    s = slice(2, -2)
    divb = ((bx[s, s, s] - bx[1:-3, s, s]) + (by[s, s, s] - by[s, 1:-3, s]) + (bz[s, s, s] - bz[s, s, 1:-3]))
    bmax = np.abs(uf[0:3]).max()
    assert bmax > 1e-6
    assert np.abs(divb).max() <= 1e-12 * bmax
    assert np.abs(G0).max() > 1e-3
    assert np.abs(G1 - G0).max() <= 1e-11 * np.abs(G0).max()
    ex, ey, ez = uf[3], uf[4], uf[5]
    dive = (ex[3:-1, s, s] - ex[s, s, s]) + (ey[s, 3:-1, s] - ey[s, s, s]) + (ez[s, s, 3:-1] - ez[s, s, s])
    assert np.abs(dive).max() > 1e3 * np.abs(G1 - G0).max()


def test_cg_stop_after_ite_max():
    """cgm stops the run after 100 iterations when the tolerance cannot be
    met (field.f90:427-430): the oracle reports the reference's message"""
    p, k, K, B, E, J = _mode2d((1, 2), seed=3)
    shape, lo = _grid2d(p)
    uf = fields_of(B, E, k[:2], OFF2, shape, lo)
    o = po.Oracle2D(p)
    po.set_cg_err(1e-300)
    with pytest.raises(RuntimeError, match="stop at cgm after ite_max"):
        o.field_from_current(uf)
