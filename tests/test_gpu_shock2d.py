"""The shock boundaries on the device (2d/proj/shock/boundary_shock.f90) through the C-ABI:
boundary_shock__injection and __particle_x bit-exact against the oracle, whole steps of the shock
app's loop order (2d/proj/shock/app.f90:111-116) bit-exact in parity mode and within the tolerance
written in the test in fast mode, with the box fully expanded and with its active edge nxe below
nxge; inject and relocate (app.f90:615-856) against the reference's formulas, then steps against
the oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po
from shock_state import shock_state

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pic(p, mode):
    from wumingpic_amd.pic2d import Pic2D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    return Pic2D(p.nx, p.ny, q, r, nsp=p.nsp, delx=p.delx, delt=p.delt, c=p.c, gfac=p.gfac, bc=po.BC_SHOCK,
                 mode=mode, capacity=p.np * p.ny)


def _active_equal(a, b, np2):
    for s in range(np2.shape[1]):
        for j in range(np2.shape[0]):
            n = np2[j, s]
            if not np.array_equal(_bits(a[:, :n, j, s]), _bits(b[:, :n, j, s])):
                return False, (s, j)
    return True, None


@pytest.mark.parametrize("which", ["injection", "particle_x"])
def test_x_boundaries_match_oracle(which):
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = shock_state(40, 8, 8, 7, u0=-0.35)
    rng = np.random.default_rng(8)
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            up[0, :n, j, s] = p.nxgs + rng.random(n) * (p.nx + 0.9)
    print(f"shock x boundary {which}", flush=True)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_gp(up, np2)
        if which == "injection":
            sim.bc_injection(p.u0)
        else:
            sim.bc_particle_x()
        gp, n2 = sim.download_gp(p.np)
    ref = up.copy(order="F")
    o = po.Oracle2D(p)
    if which == "injection":
        o.bc_injection(ref, np2)
    else:
        o.bc_particle_x(ref, np2)
    assert np.array_equal(n2, np2)
    ok, info = _active_equal(gp, ref, np2)
    assert ok, info


def test_shock_steps_parity_mode_bit_exact():
    from wumingpic_amd import MODE_PARITY
    p, up, uf, cc, np2 = shock_state(48, 16, 12, 9, u0=-0.35)
    print("shock steps parity", flush=True)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(p.u0)
        sim.step(12)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 12)
    assert np.array_equal(n2_dev, np2) and np.array_equal(cc_dev, cc)
    ok, info = _active_equal(u_dev, up, np2)
    assert ok, info
    assert np.array_equal(_bits(uf_dev), _bits(uf))


def test_shock_steps_fast_mode_close():
    """fast mode reorders the deposit's and CG's sums: fields within 1e-9 of their maximum"""
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = shock_state(48, 16, 12, 10, u0=-0.35)
    n0 = np2.sum(axis=0)
    print("shock steps fast", flush=True)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(p.u0)
        sim.step(12)
        uf_dev = sim.download_fields()
        counts = sim.counts()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 12)
    assert list(counts) == list(n0)
    assert np.abs(uf_dev - uf).max() <= 1e-9 * np.abs(uf).max()


def test_shock_rejected_in_the_mixed_form():
    from wumingpic_amd import MODE_MIXED, WpicError
    from wumingpic_amd.pic2d import Pic2D
    with pytest.raises(WpicError, match="FP64"):
        Pic2D(32, 16, [0.1, -0.1], [25.0, 1.0], bc=po.BC_SHOCK, mode=MODE_MIXED, capacity=1024)


def test_inject_then_steps_match_oracle():
    """inject (2d/proj/shock/app.f90:697-856) between parity steps: the sorted rows keep their
    particles in order and gain nl[r] particles at their end with the reference's positions, ids and
    row membership; the upstream field is set at nxe-1 and nxe; the steps that follow equal the
    oracle's from the injected state bit for bit"""
    from wumingpic_amd import MODE_PARITY
    p, up, uf, cc, np2 = shock_state(48, 16, 12, 11, u0=-0.35)
    u0 = p.u0
    v0 = u0 / np.sqrt(1 + u0 * u0)
    b0, theta, phi, l_damp = 0.5 / abs(p.q[0]), np.pi / 3, np.pi / 2, 10.0
    nl = np.random.default_rng(12).integers(2, 7, size=p.ny).astype(np.int32)
    print("shock inject", flush=True)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(u0)
        sim.step(2)
        u_a, n_a, c_a = sim.download_up(p.np)
        sim.inject(5, 2, nl, v0, 0.02, 0.1, l_damp, b0, theta, phi)
        u_b, n_b, c_b = sim.download_up(p.np)
        f_b = sim.download_fields()
        sim.step(4)
        u_c, n_c, c_c = sim.download_up(p.np)
        f_c = sim.download_fields()
    # the injection
    assert np.array_equal(n_b, n_a + nl[:, None])
    for s in range(2):
        tot = int(n_a[:, s].sum())
        for j in range(p.ny):
            n = n_a[j, s]
            assert np.array_equal(_bits(u_b[:, :n, j, s]), _bits(u_a[:, :n, j, s]))
            new = u_b[:, n:n + nl[j], j, s]
            m = np.arange(1, nl[j] + 1)
            ids = new[5].view(np.int64)
            assert np.array_equal(ids, -(m + tot + nl[:j].sum()))
            assert np.all(np.floor(new[1]) == p.nys + j)
            x_lat = p.nxge + (m - 0.5) / nl[j] * abs(v0) * p.delt + v0 * p.delt
            assert np.all(np.abs(new[0] - x_lat) < 6 * (0.02, 0.1)[s] * p.delt)
        assert np.array_equal(u_b[1, n_a[0, 0]:n_b[0, 0], 0, 0], u_b[1, n_a[0, 1]:n_b[0, 1], 0, 1])
    assert np.array_equal(c_b[:-2], c_a[:-2])
    assert np.array_equal(c_b[-2:], c_a[-2:] + nl[None, :, None])
    by, bz = b0 * np.sin(theta) * np.cos(phi), b0 * np.sin(theta) * np.sin(phi)
    ie = p.nxge - (p.nxgs - 2)
    assert np.all(f_b[1, ie - 1:ie + 1, :] == by) and np.all(f_b[2, ie - 1:ie + 1, :] == bz)
    assert np.all(f_b[4, ie - 1, :] == v0 * bz) and np.all(f_b[5, ie - 1, :] == -v0 * by)
    # the oracle: the same first two steps, then the injected state
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 2)
    ok, info = _active_equal(u_a, up, np2)
    assert ok, info
    up2, uf2 = u_b.copy(order="F"), f_b.copy(order="F")
    cc2, np22 = c_b.copy(order="F"), n_b.copy(order="F")
    o.step(up2, gp, uf2, cc2, np22, 4)
    assert np.array_equal(n_c, np22) and np.array_equal(c_c, cc2)
    ok, info = _active_equal(u_c, up2, np22)
    assert ok, info
    assert np.array_equal(_bits(f_c), _bits(uf2))


@pytest.mark.parametrize("mode", ["fast", "parity"])
def test_reflections_at_the_injection_edge_match_oracle(mode):
    """particles just left of the moving mirror with ux > u0 reflect to 2 u0 - ux there, which
    changes their gamma: the deposit's vz comes from the reflected momenta (field.f90:270-272).
    Parity mode bit-exact, fast mode fields within 1e-9 of their maximum."""
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    import states2d as S
    nx, ny, n_ppc = 24, 8, 8
    u0 = -0.35
    rng = np.random.default_rng(21)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r, bc=po.BC_SHOCK, u0=u0)
    counts = np.zeros((nx, ny), dtype=np.int64)
    counts[nx - 4:nx - 1, :] = n_ppc  # the three cells next to the right edge
    counts[2:6, :] = n_ppc // 2
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.1, 0.3), drift=(0.0, 0.0), colocate=True,
                                     counts=[counts, counts])
    for s in range(2):
        for j in range(ny):
            up[2, :np2[j, s], j, s] = np.abs(up[2, :np2[j, s], j, s]) + 0.2  # moving right, ux > u0
    uf = S.periodic_fill_fields(p, S.random_fields(p, rng, scale=0.01))
    print(f"shock reflections {mode}", flush=True)
    with _pic(p, MODE_FAST if mode == "fast" else MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(u0)
        sim.step(4)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 4)
    # many did reflect: only the mirror turns an electron's ux negative here
    nref = sum(int((up[2, :np2[j, 1], j, 1] < 0).sum()) for j in range(ny))
    assert nref > 0.1 * np2[:, 1].sum(), nref
    if mode == "parity":
        assert np.array_equal(n2_dev, np2) and np.array_equal(cc_dev, cc)
        ok, info = _active_equal(u_dev, up, np2)
        assert ok, info
        assert np.array_equal(_bits(uf_dev), _bits(uf))
    else:
        assert np.abs(uf_dev - uf).max() <= 1e-9 * np.abs(uf).max()


@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_active_edge_steps_match_oracle(mode):
    """the expanding box (nxe < nxge, 2d/proj/shock/app.f90:276-277): steps with the active edge
    set on the device equal the oracle's with the same edge, bit for bit in parity mode (particles,
    cell offsets up to nxe+1, every field value, the ones beyond nxe+2 untouched), within 1e-9 of
    the field's maximum in fast mode"""
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    from shock_state import wider_box
    p0, up, uf0, cc0, np2 = shock_state(40, 16, 10, 21, u0=-0.35)
    pb, uf, cc = wider_box(p0, uf0, cc0, 24, np.random.default_rng(22))
    print(f"shock active edge {mode}: nxe {pb.nxe} of nxge {pb.nxge}", flush=True)
    with _pic(pb, MODE_PARITY if mode == "parity" else MODE_FAST) as sim:
        sim.nxe = pb.nxe
        assert sim.nxe == pb.nxe
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(pb.u0)
        sim.step(10)
        u_dev, n2_dev, cc_dev = sim.download_up(pb.np)
        uf_dev = sim.download_fields()
    beyond = uf[:, p0.nx + 4:, :].copy()
    o = po.Oracle2D(pb)
    o.step(up, np.zeros_like(up, order="F"), uf, cc, np2, 10)
    assert np.array_equal(n2_dev, np2)
    assert np.array_equal(uf_dev[:, p0.nx + 4:, :], beyond)
    if mode == "parity":
        assert np.array_equal(cc_dev[:p0.nx + 1], cc[:p0.nx + 1])
        ok, info = _active_equal(u_dev, up, np2)
        assert ok, info
        assert np.array_equal(_bits(uf_dev), _bits(uf))
    else:
        assert np.abs(uf_dev - uf).max() <= 1e-9 * np.abs(uf).max()


def test_relocate_then_steps_match_oracle():
    """relocate (2d/proj/shock/app.f90:615-692) after inject, as the app's loop calls them
    (:120-125): nxe grows by one; each row of each species keeps its particles in order and gains
    n0 at its end on the lattice (nxe-1) dx + (m-1/2)/n0 dx with y in the row (shared by the
    species) and ids -(m + r n0 + nptotal); cumcnt(nxe) = cumcnt(nxe-1) + n0; the upstream field is
    set at nxe-1 and nxe; the parity steps that follow equal the oracle's from the relocated
    state, bit for bit; at nxe = nxge relocate does nothing"""
    from wumingpic_amd import MODE_PARITY
    from shock_state import wider_box
    p0, up, uf0, cc0, np2 = shock_state(40, 16, 10, 31, u0=-0.35)
    pb, uf, cc = wider_box(p0, uf0, cc0, 3, np.random.default_rng(32))
    u0 = pb.u0
    v0 = u0 / np.sqrt(1 + u0 * u0)
    b0, theta, phi, l_damp, n0 = 0.5 / abs(pb.q[0]), np.pi / 3, np.pi / 2, 10.0, 6
    nl = np.random.default_rng(33).integers(2, 7, size=pb.ny).astype(np.int32)
    by, bz = b0 * np.sin(theta) * np.cos(phi), b0 * np.sin(theta) * np.sin(phi)
    print(f"shock relocate: nxe {pb.nxe} of nxge {pb.nxge}", flush=True)
    with _pic(pb, MODE_PARITY) as sim:
        sim.nxe = pb.nxe
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(u0)
        sim.step(2)
        sim.inject(5, 2, nl, v0, 0.02, 0.1, l_damp, b0, theta, phi)
        u_a, n_a, c_a = sim.download_up(pb.np)
        f_a = sim.download_fields()
        sim.relocate(5, 2, n0, v0, 0.02, 0.1, l_damp, b0, theta, phi)
        nxe = sim.nxe
        u_b, n_b, c_b = sim.download_up(pb.np)
        f_b = sim.download_fields()
        sim.step(4)
        u_c, n_c, c_c = sim.download_up(pb.np)
        f_c = sim.download_fields()
        sim.relocate(5, 6, n0, v0, 0.02, 0.1, l_damp, b0, theta, phi)
        sim.relocate(5, 6, n0, v0, 0.02, 0.1, l_damp, b0, theta, phi)
        assert sim.nxe == pb.nxge
        u_d, n_d, c_d = sim.download_up(pb.np)
        sim.relocate(5, 7, n0, v0, 0.02, 0.1, l_damp, b0, theta, phi)
        assert sim.nxe == pb.nxge
        u_e, n_e, c_e = sim.download_up(pb.np)
    assert nxe == pb.nxe + 1
    assert np.array_equal(n_b, n_a + n0)
    ie = nxe - pb.nxgs  # cumcnt(nxe)
    for s in range(2):
        tot = int(n_a[:, s].sum())
        for j in range(pb.ny):
            n = n_a[j, s]
            assert np.array_equal(_bits(u_b[:, :n, j, s]), _bits(u_a[:, :n, j, s]))
            new = u_b[:, n:n + n0, j, s]
            m = np.arange(1, n0 + 1)
            assert np.array_equal(new[0], (nxe - 1) * 1.0 + (m - 0.5) / n0 * 1.0)
            assert np.all(np.floor(new[1]) == pb.nys + j)
            assert np.array_equal(new[5].view(np.int64), -(m + j * n0 + tot))
            assert np.array_equal(c_b[:ie, j, s], c_a[:ie, j, s])
            assert c_b[ie, j, s] == c_a[ie - 1, j, s] + n0
            assert np.all(c_b[ie + 1:, j, s] == n_b[j, s])
        assert np.array_equal(u_b[1, n_a[0, 0]:n_b[0, 0], 0, 0], u_b[1, n_a[0, 1]:n_b[0, 1], 0, 1])
    fi = nxe - (pb.nxgs - 2)
    assert np.all(f_b[1, fi - 1:fi + 1, :] == by) and np.all(f_b[2, fi - 1:fi + 1, :] == bz)
    assert np.all(f_b[4, fi - 1, :] == v0 * bz) and np.all(f_b[5, fi - 1, :] == -v0 * by)
    changed = np.zeros(f_b.shape[1], dtype=bool)
    changed[fi - 1:fi + 1] = True
    assert np.array_equal(_bits(f_b[:, ~changed, :]), _bits(f_a[:, ~changed, :]))
    # the oracle: the same first two steps (its field solver's state then equals the device's),
    # then the relocated state with the grown edge
    o = po.Oracle2D(pb)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 2)
    pb.nxe = nxe
    up2, uf2, cc2, np22 = u_b.copy(order="F"), f_b.copy(order="F"), c_b.copy(order="F"), n_b.copy(order="F")
    o.step(up2, gp, uf2, cc2, np22, 4)
    assert np.array_equal(n_c, np22) and np.array_equal(c_c[:ie + 2], cc2[:ie + 2])
    ok, info = _active_equal(u_c, up2, np22)
    assert ok, info
    assert np.array_equal(_bits(f_c), _bits(uf2))
    # two more relocations reach nxge; a third changes nothing
    assert np.array_equal(n_d, n_c + 2 * n0)
    assert np.array_equal(n_e, n_d) and np.array_equal(c_e, c_d)
    for s in range(2):
        for j in range(pb.ny):
            assert np.array_equal(_bits(u_e[:, :n_e[j, s], j, s]), _bits(u_d[:, :n_d[j, s], j, s]))
