"""The Vay kick on the device (particle__solv_vay, 2d/common/particle.f90:182-315) through the
C-ABI: bit-exact against the reference's own routine (golden vectors) and against the oracle on
fresh states, in both arithmetic modes; whole steps with the Vay kick in parity mode equal the
oracle's loop built from the same routines bit for bit, and fast mode stays within the tolerance
written in the test."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pic(p, mode, capacity=None):
    from wumingpic_amd.pic2d import Pic2D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    return Pic2D(p.nx, p.ny, q, r, nsp=p.nsp, delx=p.delx, delt=p.delt, c=p.c, gfac=p.gfac, bc=p.bc, mode=mode,
                 capacity=capacity or p.np * p.ny)


def _active_equal(a, b, np2):
    for s in range(np2.shape[1]):
        for j in range(np2.shape[0]):
            n = np2[j, s]
            if not np.array_equal(_bits(a[:, :n, j, s]), _bits(b[:, :n, j, s])):
                return False, (s, j)
    return True, None


@pytest.mark.parametrize("mode", ["fast", "parity"])
def test_vay_matches_reference_golden(mode):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    d = np.load(os.path.join(GOLD, "ref2d_vay.npz"))
    nx, ny = int(d["nxge"] - d["nxgs"] + 1), int(d["nye"] - d["nys"] + 1)
    p = po.make_params(nx, ny, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])
    up, uf, cc, np2 = (np.asfortranarray(d[k]) for k in ("up", "uf", "cumcnt", "np2"))
    print(f"vay golden {mode}", flush=True)
    with _pic(p, MODE_FAST if mode == "fast" else MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.particle_solv_vay()
        gp, n2 = sim.download_gp(p.np)
        # the standalone call leaves the context's kick as it was (Buneman-Boris)
        sim.particle_solv()
        gb, _ = sim.download_gp(p.np)
    assert np.array_equal(n2, np2)
    ok, info = _active_equal(gp, d["gp"], np2)
    assert ok, info
    gbo = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv(gbo, up, uf, cc)
    ok, info = _active_equal(gb, gbo, np2)
    assert ok, info


@pytest.mark.parametrize("seed", [11, 12])
def test_vay_fresh_states_match_oracle(seed):
    from wumingpic_amd import MODE_FAST
    rng = np.random.default_rng(seed)
    nx, ny, n_ppc = 40, 12, 10
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(0, 2 * n_ppc, size=(nx, ny)) for _ in range(2)]
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                     counts=counts)
    uf = S.random_fields(p, rng, scale=3.0 / abs(q[1]), b0=1.0 / abs(q[1]))
    print(f"vay fresh seed {seed}", flush=True)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.particle_solv_vay()
        gp, _ = sim.download_gp(p.np)
    ref = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv_vay(ref, up, uf, cc)
    ok, info = _active_equal(gp, ref, np2)
    assert ok, info


def _oracle_vay_steps(p, up, uf, cc, np2, nsteps):
    """the Weibel loop order (2d/proj/weibel/app.f90:101-106) with particle__solv_vay"""
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    for _ in range(nsteps):
        o.particle_solv_vay(gp, up, uf, cc)
        o.field_fdtd_i(uf, up, gp, cc)
        o.bc_particle_x(gp, np2)
        o.bc_particle_y(gp, np2)
        o.sort_bucket(up, gp, cc, np2)


def _beams(mode, nx=32, ny=32, n_ppc=16):
    from wumingpic_amd.pic2d import weibel_beams_2d
    sim = weibel_beams_2d(nx, ny, n_ppc, seed=77, mode=mode, capacity=2 * n_ppc * nx * ny)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, np2, cc = sim.download_up(p.np)
    uf = sim.download_fields()
    return sim, p, up, uf, cc, np2


def test_vay_steps_parity_mode_bit_exact():
    from wumingpic_amd import MODE_PARITY
    from wumingpic_amd._lib import PUSHER_VAY
    sim, p, up, uf, cc, np2 = _beams(MODE_PARITY)
    print("vay steps parity", flush=True)
    with sim:
        sim.set_pusher(PUSHER_VAY)
        sim.step(8)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    _oracle_vay_steps(p, up, uf, cc, np2, 8)
    assert np.array_equal(n2_dev, np2) and np.array_equal(cc_dev, cc)
    ok, info = _active_equal(u_dev, up, np2)
    assert ok, info
    assert np.array_equal(_bits(uf_dev), _bits(uf))


def test_vay_steps_fast_mode_close():
    """fast mode reorders the deposit's sums: fields within 1e-9 of their maximum after 8 steps"""
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd._lib import PUSHER_VAY
    sim, p, up, uf, cc, np2 = _beams(MODE_FAST)
    print("vay steps fast", flush=True)
    with sim:
        sim.set_pusher(PUSHER_VAY)
        sim.step(8)
        uf_dev = sim.download_fields()
        counts = sim.counts()
    _oracle_vay_steps(p, up, uf, cc, np2, 8)
    assert list(counts) == [16 * 32 * 32] * 2
    assert np.abs(uf_dev - uf).max() <= 1e-9 * np.abs(uf).max()


def test_vay_rejected_in_the_mixed_form():
    from wumingpic_amd import MODE_MIXED, WpicError
    from wumingpic_amd._lib import PUSHER_BORIS, PUSHER_VAY
    p = po.make_params(32, 16, np_cap=64 * 32, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    with _pic(p, MODE_MIXED, capacity=1024) as sim:
        with pytest.raises(WpicError, match="FP64"):
            sim.set_pusher(PUSHER_VAY)
        sim.set_pusher(PUSHER_BORIS)


# ---------------------------------------------------------------------------
# 3D (3d/common/particle.f90:236-419)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["fast", "parity"])
def test_vay3d_matches_reference_golden(mode):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    from wumingpic_amd.pic3d import Pic3D
    d = np.load(os.path.join(GOLD, "ref3d_vay.npz"))
    nx, ny, nz = (int(d[f"n{a}ge"] - d[f"n{a}gs"] + 1) for a in "xyz")
    npc = int(d["np"])
    up, uf, cc, np2 = (np.asfortranarray(d[k]) for k in ("up", "uf", "cumcnt", "np2"))
    print(f"vay3d golden {mode}", flush=True)
    with Pic3D(nx, ny, nz, list(d["q"]), list(d["r"]), delt=float(d["delt"]),
               mode=MODE_FAST if mode == "fast" else MODE_PARITY, capacity=npc * ny * nz) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.particle_solv_vay()
        gp, n2 = sim.download_gp(npc)
    assert np.array_equal(n2, np2)
    for s in range(2):
        for k in range(nz):
            for j in range(ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(gp[:, :n, j, k, s]), _bits(d["gp"][:, :n, j, k, s])), (s, j, k)


def test_vay3d_steps_parity_mode_bit_exact():
    """the 3D Weibel loop order (3d/proj/weibel/app.f90:99-108) with the Vay kick, 4 steps"""
    from wumingpic_amd import MODE_PARITY
    from wumingpic_amd._lib import PUSHER_VAY
    from wumingpic_amd.pic3d import weibel_beams_3d
    nx, ny, nz, n_ppc = 16, 8, 8, 4
    sim = weibel_beams_3d(nx, ny, nz, n_ppc, seed=31, mode=MODE_PARITY, capacity=2 * n_ppc * nx * ny * nz)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    print("vay3d steps parity", flush=True)
    with sim:
        up, np2, cc = sim.download_up(p.np)
        uf = sim.download_fields()
        sim.set_pusher(PUSHER_VAY)
        sim.step(4)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    for _ in range(4):
        o.particle_solv_vay(gp, up, uf, cc)
        o.field_fdtd_i(uf, up, gp, cc)
        o.bc_particle_x(gp, np2)
        o.bc_particle_yz(gp, np2)
        o.sort_bucket(up, gp, cc, np2)
    assert np.array_equal(n2_dev, np2) and np.array_equal(cc_dev, cc)
    for s in range(2):
        for k in range(nz):
            for j in range(ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(u_dev[:, :n, j, k, s]), _bits(up[:, :n, j, k, s])), (s, j, k)
    assert np.array_equal(_bits(uf_dev), _bits(uf))
