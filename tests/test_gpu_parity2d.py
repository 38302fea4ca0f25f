"""GPU parity of the 2D step loop through the C-ABI, against the oracle and
the reference's own golden vectors.

Bar: bit-exact for the push, particle boundaries, exchange and sort (and for
the whole step in parity mode); fast mode (atomics in the deposit, tree dot
products in CG) within the tolerances written in each test.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S

pytestmark = pytest.mark.gpu

GOLD = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def _pic(p, mode, bc=None, capacity=None):
    from wumingpic_amd.pic2d import Pic2D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    cap = capacity or p.np * p.ny
    return Pic2D(p.nx, p.ny, q, r, nsp=p.nsp, delx=p.delx, delt=p.delt, c=p.c, gfac=p.gfac,
                 bc=p.bc if bc is None else bc, mode=mode, capacity=cap)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _active_equal(a, b, np2):
    """bit equality of the active slots of two up/gp arrays"""
    for s in range(np2.shape[1]):
        for j in range(np2.shape[0]):
            n = np2[j, s]
            if not np.array_equal(_bits(a[:, :n, j, s]), _bits(b[:, :n, j, s])):
                bad = np.nonzero((_bits(a[:, :n, j, s]) != _bits(b[:, :n, j, s])).any(axis=0))[0]
                return False, (s, j, bad[:5], a[:, bad[0], j, s], b[:, bad[0], j, s])
    return True, None


def _cells_equal(a, b, np2, cc):
    """bit equality of each cell's particles as a set (ordered by id): the
    fused fast step keeps the cells' contents, not their order within a cell"""
    for s in range(np2.shape[1]):
        for j in range(np2.shape[0]):
            for i in range(cc.shape[0] - 1):
                lo, hi = cc[i, j, s], cc[i + 1, j, s]
                if hi == lo:
                    continue
                ka = np.argsort(a[5, lo:hi, j, s].view(np.int64), kind="stable")
                kb = np.argsort(b[5, lo:hi, j, s].view(np.int64), kind="stable")
                xa, xb = _bits(a[:, lo:hi, j, s][:, ka]), _bits(b[:, lo:hi, j, s][:, kb])
                if not np.array_equal(xa, xb):
                    return False, (s, j, i)
    return True, None


def _params_from(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    return po.make_params(nx, ny, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def test_push_matches_reference_golden():
    """particle__solv on the device == the reference's particle__solv output"""
    from wumingpic_amd import MODE_FAST
    d = np.load(f"{GOLD}/ref2d_push.npz")
    p = _params_from(d)
    up, uf, cc, np2 = (np.asfortranarray(d[k]) for k in ("up", "uf", "cumcnt", "np2"))
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.particle_solv()
        gp, n2 = sim.download_gp(p.np)
    assert np.array_equal(n2, np2)
    ok, info = _active_equal(gp, d["gp"], np2)
    assert ok, info


def test_sort_matches_reference_golden():
    """sort__bucket of an unsorted gp == the reference's sort__bucket output"""
    from wumingpic_amd import MODE_FAST
    d = np.load(f"{GOLD}/ref2d_sort.npz")
    p = _params_from(d)
    up, np2 = np.asfortranarray(d["up"]), np.asfortranarray(d["np2"])
    with _pic(p, MODE_FAST) as sim:
        sim.upload_gp(up, np2)
        sim.sort_bucket()
        out, n2, cc = sim.download_up(p.np)
    assert np.array_equal(cc, d["cumcnt"])
    ok, info = _active_equal(out, d["gp"], np2)
    assert ok, info


def _physical_state(nx, ny, n_ppc, seed, bc=po.BC_PERIODIC, counts=None, fscale=0.02, u_scale=(0.02, 0.2)):
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r, bc=bc)
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=u_scale, drift=(0.0, 0.5), colocate=True,
                                     counts=counts)
    if bc == po.BC_WALLX:
        # keep particles off the wall cells, as the reconnection loading does
        for s in range(2):
            for j in range(ny):
                n = np2[j, s]
                x = up[0, :n, j, s]
                up[0, :n, j, s] = np.clip(x, p.nxgs + 1.0, p.nxge - 1.0 - 1e-9)
        up, cc = _resort(p, up, np2)
    uf = S.periodic_fill_fields(p, S.random_fields(p, rng, scale=fscale, b0=0.0))
    return p, up, uf, cc, np2


def _resort(p, up, np2):
    o = po.Oracle2D(p)
    out = np.zeros_like(up, order="F")
    cc = np.zeros((p.nx + 1, p.ny, p.nsp), dtype=np.int32, order="F")
    o.sort_bucket(out, up, cc, np2)
    return out, cc


@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_field_fdtd_i_matches_oracle(mode):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    p, up, uf, cc, np2 = _physical_state(24, 20, 16, 11)
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.particle_solv(gp, up, uf, cc)
    uf_ref = uf.copy(order="F")
    it_ref = o.field_fdtd_i(uf_ref, up, gp, cc)
    with _pic(p, MODE_PARITY if mode == "parity" else MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.particle_solv()
        it = sim.field_fdtd_i()
        uj, df, uf_dev = sim.download_current(), sim.download_dfield(), sim.download_fields()
    assert it == it_ref
    if mode == "parity":
        assert np.array_equal(_bits(uj), _bits(o.uj))
        assert np.array_equal(_bits(df), _bits(o.df))
        assert np.array_equal(_bits(uf_dev), _bits(uf_ref))
    else:
        # atomics reorder the deposit sums: relative to the largest entry
        scale = np.abs(o.uj).max()
        assert np.abs(uj - o.uj).max() <= 1e-13 * scale
        assert np.abs(uf_dev - uf_ref).max() <= 1e-11 * np.abs(uf_ref).max()


@pytest.mark.parametrize("bc", [po.BC_PERIODIC, po.BC_WALLX])
def test_boundaries_exchange_sort_match_oracle(bc):
    """bc__particle_x, bc__particle_y (row exchange order) and sort__bucket,
    bit for bit, on a gp whose particles cross rows and the periodic seam"""
    from wumingpic_amd import MODE_FAST
    rng = np.random.default_rng(5)
    p, up, uf, cc, np2 = _physical_state(16, 12, 8, 3, bc=bc)
    gp = up.copy(order="F")
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            gp[0, :n, j, s] += rng.uniform(-0.95, 0.95, n)
            gp[1, :n, j, s] += rng.uniform(-0.95, 0.95, n) * (rng.random(n) < 0.4)
    # exact edge: a particle just below the lower x boundary
    gp[0, 0, 0, 0] = np.nextafter(float(p.nxgs), 0.0)
    if bc == po.BC_WALLX:
        for s in range(2):
            for j in range(p.ny):
                n = np2[j, s]
                gp[0, :n, j, s] = np.clip(gp[0, :n, j, s], p.nxgs + 0.05, p.nxge + 0.95)
    o = po.Oracle2D(p)
    g_ref = gp.copy(order="F")
    n2_ref = np2.copy(order="F")
    o.bc_particle_x(g_ref, n2_ref)
    o.bc_particle_y(g_ref, n2_ref)
    up_ref = np.zeros_like(up, order="F")
    cc_ref = np.zeros_like(cc, order="F")
    o.sort_bucket(up_ref, g_ref, cc_ref, n2_ref)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_gp(gp, np2)
        sim.bc_particle_x()
        sim.bc_particle_y()
        g_dev, n2_dev = sim.download_gp(p.np)
        assert np.array_equal(n2_dev, n2_ref)
        ok, info = _active_equal(g_dev, g_ref, n2_ref)
        assert ok, info
        sim.sort_bucket()
        u_dev, n2b, cc_dev = sim.download_up(p.np)
    assert np.array_equal(cc_dev, cc_ref)
    ok, info = _active_equal(u_dev, up_ref, n2_ref)
    assert ok, info


def _weibel_device_state(nx, ny, n_ppc, mode, seed=20261202):
    from wumingpic_amd.pic2d import weibel_beams_2d
    sim = weibel_beams_2d(nx, ny, n_ppc, seed=seed, mode=mode, capacity=2 * n_ppc * nx * ny)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, np2, cc = sim.download_up(p.np)
    uf = sim.download_fields()
    return sim, p, up, uf, cc, np2


@pytest.mark.parametrize("fused", [True, False])
def test_weibel_steps_parity_mode_bit_exact(fused):
    """Config-0 physics (beams, dt = 0.5, omega_pe dt = 0.05) on a 32x32 grid:
    after 12 steps the device state equals the oracle's bit for bit."""
    from wumingpic_amd import MODE_PARITY
    sim, p, up, uf, cc, np2 = _weibel_device_state(32, 32, 16, MODE_PARITY)
    with sim:
        nsteps = 12
        if fused:
            sim.step(nsteps)
        else:
            for _ in range(nsteps):
                sim.particle_solv()
                sim.field_fdtd_i()
                sim.bc_particle_x()
                sim.bc_particle_y()
                sim.sort_bucket()
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, nsteps)
    assert np.array_equal(n2_dev, np2)
    assert np.array_equal(cc_dev, cc)
    ok, info = _active_equal(u_dev, up, np2)
    assert ok, info
    assert np.array_equal(_bits(uf_dev), _bits(uf))


def test_weibel_steps_fast_mode_close_and_conserving():
    """Fast mode on the same workload: counts and cells track the oracle, the
    fields agree to a small relative tolerance, and Gauss's law holds."""
    from wumingpic_amd import MODE_FAST
    sim, p, up, uf, cc, np2 = _weibel_device_state(32, 32, 16, MODE_FAST)
    with sim:
        G0 = sim.gauss()
        sim.step(12)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
        G1 = sim.gauss()
        counts = sim.counts()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 12)
    assert list(counts) == [16 * 32 * 32] * 2
    scale = np.abs(uf).max()
    assert np.abs(uf_dev - uf).max() <= 1e-9 * scale
    # Gauss residual stays at round-off relative to 4 pi |rho_species|
    rho_scale = 4 * np.pi * abs(p.q[1]) * 16
    assert np.abs(G1).max() <= 1e-12 * rho_scale + np.abs(G0).max() * 4


def test_walls_steps_parity_mode_bit_exact():
    """Reflecting x walls (reconnection boundary order) in parity mode"""
    from wumingpic_amd import MODE_PARITY
    p, up, uf, cc, np2 = _physical_state(24, 16, 8, 21, bc=po.BC_WALLX, fscale=0.0)
    with _pic(p, MODE_PARITY, capacity=3 * 8 * 24 * 16) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.step(10)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 10)
    assert np.array_equal(n2_dev, np2)
    assert np.array_equal(cc_dev, cc)
    ok, info = _active_equal(u_dev, up, np2)
    assert ok, info
    assert np.array_equal(_bits(uf_dev), _bits(uf))


@pytest.mark.parametrize("bc", [po.BC_PERIODIC, po.BC_WALLX])
def test_fast_mode_one_step_particles_exact_fields_close(bc):
    """Fast mode's fused step: the push, walls and wraps are the parity
    kernels' arithmetic and the exchange and sort put every particle in the
    reference's cell, so after one step each cell holds the oracle's
    particles bit for bit (in an order of the fast sort's own) with the
    oracle's cumcnt and np2; the factored deposit changes only the rounding of
    uj, so the fields agree to a relative 1e-12."""
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = _physical_state(24, 16, 8, 23, bc=bc, fscale=0.0 if bc == po.BC_WALLX else 0.02)
    with _pic(p, MODE_FAST, capacity=3 * 8 * 24 * 16) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.step(1)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
        uj_dev = sim.download_current()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 1)
    assert np.array_equal(n2_dev, np2)
    assert np.array_equal(cc_dev, cc)
    ok, info = _cells_equal(u_dev, up, np2, cc)
    assert ok, info
    assert np.abs(uj_dev - o.uj).max() <= 1e-13 * np.abs(o.uj).max()
    assert np.abs(uf_dev - uf).max() <= 1e-12 * np.abs(uf).max()
