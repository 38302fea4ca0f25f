"""The oracle's shock boundaries (2d/proj/shock/boundary_shock.f90 with nxs = nxgs).
boundary_shock.f90 needs MPI, so no reference build pins them; these tests check what the
mathematics fixes: the injection mirror is an involution, with u0 = 0 it is the static mirror
x -> 2 nxe dx - x, u -> -u; the left wall is the reconnection's; after a field solve the field
and CG ghosts beyond nxe are zero and the left ghosts are the reconnection's; with an active edge
nxe below nxge (the expanding box) the steps equal those of the box cut at nxe."""
import numpy as np

from oracle import pyoracle as po
from shock_state import shock_state


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _spread(p, np2, rng, up):
    """positions over the whole box and beyond the injection edge"""
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            up[0, :n, j, s] = p.nxgs + rng.random(n) * (p.nx + 0.9)
    return up


def test_injection_is_an_elastic_bounce_off_the_moving_mirror():
    """at small momenta (v = u to O(u^3)) a particle beyond xend comes back as far inside as it
    was beyond, with its velocity relative to the mirror (v0 = u0 / gamma0) reversed in x, and
    its y and z momenta reversed"""
    u0 = -2e-4
    p, up, uf, cc, np2 = shock_state(24, 6, 8, 1, u0=u0, vth=(1e-4, 1e-4))
    rng = np.random.default_rng(2)
    up = _spread(p, np2, rng, up)
    a = up.copy(order="F")
    po.Oracle2D(p).bc_injection(a, np2)
    v0 = u0 / np.sqrt(1 + u0 * u0)
    xend = p.nxge * p.delx + v0 * p.delt
    right = up[0] > xend
    assert right.sum() > 10
    assert np.allclose(a[0][right] - xend, -(up[0][right] - xend), rtol=0, atol=1e-12)
    vel = lambda u: u[2] / np.sqrt(1 + (u[2:5] ** 2).sum(axis=0))  # noqa: E731
    # (the residual is O(u^3) ~ 1e-11 here; a wrong sign or factor gives ~ 1e-4)
    assert np.allclose(vel(a[:, right]) - v0, -(vel(up[:, right]) - v0), rtol=0, atol=1e-9)
    assert np.array_equal(a[3:5, right], -up[3:5, right])
    # particles between the walls are left alone
    mid = (up[0] >= p.nxgs + 1) & ~right
    assert np.array_equal(_bits(a[:, mid]), _bits(up[:, mid]))


def test_injection_with_u0_zero_is_the_static_mirror_and_left_wall_is_the_reconnections():
    p, up, uf, cc, np2 = shock_state(24, 6, 8, 3, u0=0.0)
    rng = np.random.default_rng(4)
    up = _spread(p, np2, rng, up)
    a = up.copy(order="F")
    po.Oracle2D(p).bc_injection(a, np2)
    xe = p.nxge * p.delx
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            x = up[0, :n, j, s]
            r = x > xe
            assert np.array_equal(a[0, :n, j, s][r], 2 * xe - x[r])
            assert np.array_equal(a[2:5, :n, j, s][:, r], -up[2:5, :n, j, s][:, r])
    # the left wall: the reconnection's boundary on the particles left of nxgs + 1
    pw = po.make_params(p.nx, p.ny, np_cap=p.np, delt=p.delt, q=p.q, r=p.r, bc=po.BC_WALLX)
    w = up.copy(order="F")
    po.Oracle2D(pw).bc_particle_x(w, np2)
    left = up[0] < p.nxgs + 1
    assert np.array_equal(_bits(a[:, left]), _bits(w[:, left]))


def test_field_solve_ghosts():
    p, up, uf, cc, np2 = shock_state(24, 8, 8, 5)
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 3)
    i_e = p.nxge + 1 - (p.nxgs - 2)
    assert np.all(o.df[:, i_e, :] == 0.0)
    i_s = p.nxgs - 1 - (p.nxgs - 2)
    assert np.array_equal(o.df[0, i_s, :], -o.df[0, i_s + 1, :])
    assert np.array_equal(o.df[1:4, i_s, :], o.df[1:4, i_s + 2, :])
    assert np.array_equal(o.df[4:6, i_s, :], -o.df[4:6, i_s + 1, :])
    # the right edge is the only difference from the reconnection's boundaries: with the same
    # state, the walls give a different field near nxe and the same far from it after one step
    pw = po.make_params(p.nx, p.ny, np_cap=p.np, delt=p.delt, q=p.q, r=p.r, bc=po.BC_WALLX)
    _, up1, uf1, cc1, np21 = shock_state(24, 8, 8, 5)
    ufw = uf1.copy(order="F")
    ow = po.Oracle2D(pw)
    upw = up1.copy(order="F")
    ow.step(upw, np.zeros_like(up1, order="F"), ufw, cc1.copy(order="F"), np21.copy(order="F"), 1)
    ufs = uf1.copy(order="F")
    os_ = po.Oracle2D(p)
    os_.step(up1.copy(order="F"), np.zeros_like(up1, order="F"), ufs, cc1.copy(order="F"), np21.copy(order="F"), 1)
    assert not np.array_equal(ufs[:, -4:, :], ufw[:, -4:, :])
    assert np.allclose(ufs[:, 2:8, :], ufw[:, 2:8, :], rtol=0, atol=1e-3 * np.abs(ufs).max())


def test_active_edge_equals_the_box_cut_there():
    """The expanding box (nxe < nxge, 2d/proj/shock/app.f90:112-117 with nxs, nxe passed to every
    routine): steps of a box whose active edge is nxe equal, bit for bit, the steps of a box that
    ends at nxe (its particles, cell offsets and fields up to nxe+2), and leave the fields beyond
    nxe+2 untouched.  Pins the active-edge loops of the push, deposit, current fold, field solve
    (right-hand side, CG sums, E update, ghosts) and sort against the full-box code."""
    from shock_state import wider_box
    p, up, uf, cc, np2 = shock_state(24, 8, 8, 7)
    pb, ufb, ccb = wider_box(p, uf, cc, 9, np.random.default_rng(8))
    beyond = ufb[:, p.nx + 4:, :].copy()
    upb, np2b = up.copy(order="F"), np2.copy(order="F")
    o = po.Oracle2D(p)
    o.step(up, np.zeros_like(up, order="F"), uf, cc, np2, 5)
    ob = po.Oracle2D(pb)
    ob.step(upb, np.zeros_like(upb, order="F"), ufb, ccb, np2b, 5)
    assert np.array_equal(np2b, np2)
    assert np.array_equal(ccb[:p.nx + 1], cc)
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            assert np.array_equal(_bits(upb[:, :n, j, s]), _bits(up[:, :n, j, s])), (s, j)
    assert np.array_equal(_bits(ufb[:, :p.nx + 4, :]), _bits(uf))
    assert np.array_equal(_bits(ufb[:, p.nx + 4:, :]), _bits(beyond))
    assert tuple(o.st.cg_iters) == tuple(ob.st.cg_iters)
    # the same state with the whole wider box active differs (the edge is where the walls are)
    pf = po.make_params(pb.nx, p.ny, np_cap=p.np, delt=p.delt, q=[p.q[0], p.q[1]], r=[p.r[0], p.r[1]],
                        bc=p.bc, u0=p.u0)
    p2, up2, uf2, cc2, np22 = shock_state(24, 8, 8, 7)
    _, uff, ccf = wider_box(p2, uf2, cc2, 9, np.random.default_rng(8))
    po.Oracle2D(pf).step(up2, np.zeros_like(up2, order="F"), uff, ccf, np22, 1)
    assert not np.array_equal(uff[:, :p.nx + 4, :], ufb[:, :p.nx + 4, :])
