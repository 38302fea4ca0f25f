"""3D device path against the reference's own 3D particle__solv / sort__bucket
(golden vectors) and against the 3D oracle, through the C-ABI.

Parity mode is bit-exact: push, deposit (uj), field solve (df, uf, CG
iteration counts), x wrap, y-z exchange order, sort.  Fast mode reorders the
deposit's sums only: particles stay bit-exact for one step and the fields
agree to the relative tolerances written in each test.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
import states3d as S

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pic(p, mode, capacity=None):
    from wumingpic_amd.pic3d import Pic3D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    cap = capacity if capacity is not None else p.np * p.ny * p.nz
    return Pic3D(p.nx, p.ny, p.nz, q, r, nsp=p.nsp, delt=p.delt, mode=mode, capacity=cap)


def _active_equal(a, b, np2):
    ny, nz, nsp = np2.shape
    for s in range(nsp):
        for k in range(nz):
            for j in range(ny):
                n = np2[j, k, s]
                if not np.array_equal(_bits(a[:, :n, j, k, s]), _bits(b[:, :n, j, k, s])):
                    bad = np.nonzero((_bits(a[:, :n, j, k, s]) != _bits(b[:, :n, j, k, s])).any(axis=0))[0]
                    return False, (s, j, k, int(bad[0]), a[:, bad[0], j, k, s], b[:, bad[0], j, k, s])
    return True, None


def _cells_equal(a, b, np2, cc):
    """bit equality of each cell's particles as a set (ordered by id): the
    fused fast step keeps the cells' contents, not their order within a cell"""
    ny, nz, nsp = np2.shape
    for s in range(nsp):
        for k in range(nz):
            for j in range(ny):
                for i in range(cc.shape[0] - 1):
                    lo, hi = cc[i, j, k, s], cc[i + 1, j, k, s]
                    if hi == lo:
                        continue
                    ka = np.argsort(a[6, lo:hi, j, k, s].view(np.int64), kind="stable")
                    kb = np.argsort(b[6, lo:hi, j, k, s].view(np.int64), kind="stable")
                    if not np.array_equal(_bits(a[:, lo:hi, j, k, s][:, ka]), _bits(b[:, lo:hi, j, k, s][:, kb])):
                        return False, (s, j, k, i)
    return True, None


def _params(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    nz = int(d["nze"] - d["nzs"] + 1)
    return po.make_params3d(nx, ny, nz, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def test_push3d_matches_reference_golden():
    from wumingpic_amd import MODE_PARITY
    d = np.load(os.path.join(GOLD, "ref3d_push.npz"))
    p = _params(d)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_up(d["up"], d["np2"], d["cumcnt"])
        sim.upload_fields(d["uf"])
        sim.particle_solv()
        gp, n2 = sim.download_gp(p.np)
    ok, info = _active_equal(gp, d["gp"], d["np2"])
    assert ok, info


def test_sort3d_matches_reference_golden():
    from wumingpic_amd import MODE_PARITY
    d = np.load(os.path.join(GOLD, "ref3d_sort.npz"))
    p = _params(d)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_gp(d["up"], d["np2"])
        sim.sort_bucket()
        up, n2, cc = sim.download_up(p.np)
    assert np.array_equal(cc, d["cumcnt"])
    ok, info = _active_equal(up, d["gp"], d["np2"])
    assert ok, info


def _physical_state(nx, ny, nz, n_ppc, seed, fscale=0.02, u_scale=(0.02, 0.2)):
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=u_scale, drift=(0.0, 0.5), colocate=True)
    uf = S.periodic_fill_fields3d(p, S.random_fields3d(p, rng, scale=fscale))
    return p, up, uf, cc, np2


@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_field_fdtd_i3d_matches_oracle(mode):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    p, up, uf, cc, np2 = _physical_state(10, 6, 5, 4, 11)
    o = po.Oracle3D(p)
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
        # atomics reorder the deposit's sums: relative to the largest entry
        assert np.abs(uj - o.uj).max() <= 1e-13 * np.abs(o.uj).max()
        assert np.abs(uf_dev - uf_ref).max() <= 1e-11 * np.abs(uf_ref).max()


def test_boundaries_exchange_sort3d_match_oracle():
    """bc__particle_x, bc__particle_yz (the exact arrival and hole-filling
    order, across faces, edges and corners of the periodic box) and
    sort__bucket, entry point by entry point"""
    from wumingpic_amd import MODE_PARITY
    rng = np.random.default_rng(31)
    p = po.make_params3d(6, 5, 4, np_cap=60, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles3d(p, 4, rng, u_scale=(0.1, 0.1), drift=(0.0, 0.0), colocate=False)
    gp = up.copy(order="F")
    for s in range(2):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                m = rng.random(n) < 0.4
                gp[0:3, :n, j, k, s] += m * rng.uniform(-0.95, 0.95, (3, n))
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_gp(gp, np2)
        sim.bc_particle_x()
        sim.bc_particle_yz()
        gx, n2x = sim.download_gp(p.np)
        sim.sort_bucket()
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
    o = po.Oracle3D(p)
    g_ref = gp.copy(order="F")
    n2_ref = np2.copy(order="F")
    o.bc_particle_x(g_ref, n2_ref)
    o.bc_particle_yz(g_ref, n2_ref)
    assert np.array_equal(n2x, n2_ref)
    ok, info = _active_equal(gx, g_ref, n2_ref)
    assert ok, info
    u_ref = np.zeros_like(g_ref, order="F")
    cc_ref = np.zeros_like(cc, order="F")
    o.sort_bucket(u_ref, g_ref, cc_ref, n2_ref)
    assert np.array_equal(cc_dev, cc_ref)
    ok, info = _active_equal(u_dev, u_ref, n2_ref)
    assert ok, info


def _weibel3d_device_state(nx, ny, nz, n_ppc, mode, seed=20261202):
    from wumingpic_amd.pic3d import weibel_beams_3d
    sim = weibel_beams_3d(nx, ny, nz, n_ppc, seed=seed, mode=mode, capacity=2 * n_ppc * nx * ny * nz)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, np2, cc = sim.download_up(p.np)
    uf = sim.download_fields()
    return sim, p, up, uf, cc, np2


@pytest.mark.parametrize("fused", [True, False])
def test_weibel3d_steps_parity_mode_bit_exact(fused):
    """3D Weibel beams (dt = 0.5, omega_pe dt = 0.05) on 12x8x6 cells, 4 ppc:
    after 6 steps the device state equals the oracle's bit for bit."""
    from wumingpic_amd import MODE_PARITY
    sim, p, up, uf, cc, np2 = _weibel3d_device_state(12, 8, 6, 4, MODE_PARITY)
    nsteps = 6
    with sim:
        if fused:
            sim.step(nsteps)
        else:
            for _ in range(nsteps):
                sim.particle_solv()
                sim.field_fdtd_i()
                sim.bc_particle_x()
                sim.bc_particle_yz()
                sim.sort_bucket()
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, nsteps)
    assert np.array_equal(n2_dev, np2)
    assert np.array_equal(cc_dev, cc)
    ok, info = _active_equal(u_dev, up, np2)
    assert ok, info
    assert np.array_equal(_bits(uf_dev), _bits(uf))


def test_fast3d_one_step_particles_exact_fields_close():
    """Fast mode's fused step: the push and wraps are the parity kernels'
    arithmetic and the one-pass exchange and sort put every particle in the
    reference's cell, so after one step each cell holds the oracle's
    particles bit for bit (in an order of the fast sort's own) with the
    oracle's cumcnt and np2; the fields agree to a relative 1e-12."""
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = _physical_state(10, 6, 5, 4, 23)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.step(1)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 1)
    assert np.array_equal(n2_dev, np2)
    assert np.array_equal(cc_dev, cc)
    ok, info = _cells_equal(u_dev, up, np2, cc)
    assert ok, info
    assert np.abs(uf_dev - uf).max() <= 1e-12 * np.abs(uf).max()


def test_fast3d_cell_changers_beyond_the_tile_queue():
    """One 8x4x4 tile with 48 ppc per species and electrons drifting at u = 3:
    about 2,900 particles change cell in one step, more than the fused fast
    kernel queues in LDS (128 per wave and round of 8 cells, 2,048 per tile),
    so the rest go through its global overflow list.  Particles stay
    bit-exact, the current agrees to rounding."""
    from wumingpic_amd import MODE_FAST
    rng = np.random.default_rng(41)
    q, r = S.weibel_consts(48)
    p = po.make_params3d(8, 4, 4, np_cap=3 * 48 * 8, delt=0.5, q=q, r=r)
    up, cc, np2 = S.sorted_particles3d(p, 48, rng, u_scale=(0.02, 0.2), drift=(0.0, 3.0), colocate=True)
    uf = S.periodic_fill_fields3d(p, S.random_fields3d(p, rng, scale=0.02))
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.particle_solv(gp, up, uf, cc)
    uf_ref = uf.copy(order="F")
    o.field_fdtd_i(uf_ref, up, gp, cc)
    moved = 0
    for k in range(p.nz):
        for j in range(p.ny):
            n = np2[j, k, 1]
            old = np.floor(up[0:3, :n, j, k, 1] / p.delx)
            new = np.floor(gp[0:3, :n, j, k, 1] / p.delx)
            moved += int((old != new).any(axis=0).sum())
    assert moved > 2048 + 200
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.step(1)
        uj = sim.download_current()
        uf_dev = sim.download_fields()
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
    assert np.abs(uj - o.uj).max() <= 1e-13 * np.abs(o.uj).max()
    assert np.abs(uf_dev - uf_ref).max() <= 1e-12 * np.abs(uf_ref).max()
    # the particles of the whole step (the overflowed ones wrapped after their
    # deposit) equal the oracle's step bit for bit, cell by cell
    o2 = po.Oracle3D(p)
    up2, uf2, cc2, np22 = up.copy(order="F"), uf.copy(order="F"), cc.copy(order="F"), np2.copy(order="F")
    o2.step(up2, np.zeros_like(up2, order="F"), uf2, cc2, np22, 1)
    assert np.array_equal(n2_dev, np22) and np.array_equal(cc_dev, cc2)
    ok, info = _cells_equal(u_dev, up2, np22, cc2)
    assert ok, info


def test_fast3d_steps_sparse_columns_match_oracle():
    """Fast mode's one-pass exchange and sort on a ragged state: cells with 0 to
    7 particles, two empty columns, and y-z momenta large enough that many
    particles change column every step (arrivals into empty columns, columns
    that lose most of theirs, movers across every face, edge and corner of the
    periodic box).  After one step each cell holds the oracle's particles bit
    for bit as a set, with the oracle's np2 and cumcnt; two more steps (the
    sort now reading through the previous step's permutation, the push on
    fields that carry the deposit's rounding) keep every particle id exactly
    once, cumcnt consistent with np2 and the fields within 1e-9 of the
    oracle's."""
    from wumingpic_amd import MODE_FAST
    rng = np.random.default_rng(53)
    q, r = S.weibel_consts(4)
    p = po.make_params3d(6, 5, 4, np_cap=3 * 8 * 6, delt=0.5, q=q, r=r)
    counts = rng.integers(0, 8, size=(6, 5, 4))
    counts[:, 1, 2] = 0
    counts[:, 3, 0] = 0
    up, cc, np2 = S.sorted_particles3d(p, 4, rng, u_scale=(0.3, 0.6), drift=(0.0, 0.0), colocate=True,
                                       counts=[counts, counts])
    uf = S.periodic_fill_fields3d(p, S.random_fields3d(p, rng, scale=0.02))
    assert (np2 == 0).sum() == 4
    ids0 = [np.sort(np.concatenate([up[6, :np2[j, k, s], j, k, s].view(np.int64) for k in range(p.nz)
                                    for j in range(p.ny)])) for s in range(2)]
    print("fast3d sparse columns, 1 + 2 steps", flush=True)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.step(1)
        u1, n21, cc1 = sim.download_up(p.np)
        sim.step(2)
        u3, n23, cc3 = sim.download_up(p.np)
        uf3 = sim.download_fields()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 1)
    assert np.array_equal(n21, np2)
    assert np.array_equal(cc1, cc)
    ok, info = _cells_equal(u1, up, np2, cc)
    assert ok, info
    o.step(up, gp, uf, cc, np2, 2)
    for s in range(2):
        ids = np.sort(np.concatenate([u3[6, :n23[j, k, s], j, k, s].view(np.int64) for k in range(p.nz)
                                      for j in range(p.ny)]))
        assert np.array_equal(ids, ids0[s]), s
    assert np.array_equal(cc3[-1], n23)
    assert np.all(np.diff(cc3, axis=0) >= 0)
    assert np.abs(uf3 - uf).max() <= 1e-9 * np.abs(uf).max()


def test_weibel3d_fast_mode_close_and_conserving():
    from wumingpic_amd import MODE_FAST
    sim, p, up, uf, cc, np2 = _weibel3d_device_state(12, 8, 6, 4, MODE_FAST)
    with sim:
        G0 = sim.gauss()
        sim.step(8)
        uf_dev = sim.download_fields()
        G1 = sim.gauss()
        counts = sim.counts()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 8)
    assert list(counts) == [4 * 12 * 8 * 6] * 2
    assert np.abs(uf_dev - uf).max() <= 1e-9 * np.abs(uf).max()
    rho_scale = 4 * np.pi * abs(p.q[1]) * 4
    assert np.abs(G1).max() <= 1e-12 * rho_scale + np.abs(G0).max() * 4
