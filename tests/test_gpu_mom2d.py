"""GPU parity of the 2D moments (mom_calc__accl, mom_calc__nvt and the moment
boundary folds) through the C-ABI, against the reference's own golden vectors
and the oracle.

Bar: mom_calc__accl bit-exact in every mode; mom_calc__nvt and the folds
bit-exact in parity mode; fast mode (atomic summation order) within 1e-13 of
each moment's largest magnitude.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S

pytestmark = pytest.mark.gpu

GOLD = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")
FAST_TOL = 1e-13


def _pic(p, mode, bc=None):
    from wumingpic_amd.pic2d import Pic2D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    return Pic2D(p.nx, p.ny, q, r, nsp=p.nsp, delx=p.delx, delt=p.delt, c=p.c, gfac=p.gfac,
                 bc=p.bc if bc is None else bc, mode=mode, capacity=p.np * p.ny)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _close(a, b):
    """per moment component, relative to its largest magnitude"""
    for k in range(7):
        scale = max(np.abs(b[k]).max(), 1e-300)
        err = np.abs(a[k] - b[k]).max() / scale
        assert err <= FAST_TOL, (k, err)


def _golden():
    d = np.load(f"{GOLD}/ref2d_mom.npz")
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    p = po.make_params(nx, ny, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])
    return d, p


@pytest.mark.parametrize("parity", [True, False])
def test_accl_and_nvt_match_reference_golden(parity):
    """mom_calc__accl then mom_calc__nvt on the device == the reference's own"""
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    d, p = _golden()
    up, uf, cc, np2 = (np.asfortranarray(d[k]) for k in ("up", "uf", "cumcnt", "np2"))
    print(f"accl+nvt golden, parity={parity}", flush=True)
    with _pic(p, MODE_PARITY if parity else MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.mom_accl()
        gp, n2 = sim.download_gp(p.np)
        sim.mom_nvt()
        mom = sim.download_mom()
    assert np.array_equal(n2, np2)
    for s in range(p.nsp):
        for j in range(p.ny):
            n = np2[j, s]
            assert np.array_equal(_bits(gp[:5, :n, j, s]), _bits(d["gp"][:5, :n, j, s]))
            # the id travels with the record
            assert np.array_equal(_bits(gp[5, :n, j, s]), _bits(up[5, :n, j, s]))
    if parity:
        assert np.array_equal(_bits(mom), _bits(d["mom"]))
    else:
        _close(mom, d["mom"])


@pytest.mark.parametrize("bc", [po.BC_PERIODIC, po.BC_WALLX])
def test_bc_mom_matches_oracle(bc):
    """boundary_periodic__mom / boundary_reconnection__mom on an arbitrary
    array, bit for bit"""
    from wumingpic_amd import MODE_PARITY
    rng = np.random.default_rng(3 + bc)
    p = po.make_params(11, 9, np_cap=64, q=(0.1, -0.1), r=(25.0, 1.0), bc=bc)
    mom = np.asfortranarray(rng.standard_normal((7, p.nx + 2, p.ny + 2, p.nsp)))
    print(f"bc_mom bc={bc}", flush=True)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_mom(mom)
        sim.bc_mom()
        dev = sim.download_mom()
    o = po.Oracle2D(p)
    o.bc_mom(mom)
    assert np.array_equal(_bits(dev), _bits(mom))


def _weibel(nx, ny, n_ppc, mode):
    from wumingpic_amd.pic2d import weibel_beams_2d
    sim = weibel_beams_2d(nx, ny, n_ppc, mode=mode, capacity=2 * n_ppc * nx * ny)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    return sim, p


def _state(nx, ny, n_ppc, seed, bc):
    """sorted particles with beams, walls kept off the wall cells as the
    reconnection loading does, and small random fields"""
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r, bc=bc)
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.02, 0.2), drift=(0.0, 0.5), colocate=True)
    if bc == po.BC_WALLX:
        for s in range(2):
            for j in range(ny):
                n = np2[j, s]
                up[0, :n, j, s] = np.clip(up[0, :n, j, s], p.nxgs + 1.0, p.nxge - 1.0 - 1e-9)
        out = np.zeros_like(up, order="F")
        po.Oracle2D(p).sort_bucket(out, up, cc, np2)
        up = out
    uf = S.periodic_fill_fields(p, S.random_fields(p, rng, scale=0.02, b0=0.0))
    return p, up, uf, cc, np2


@pytest.mark.parametrize("bc", [po.BC_PERIODIC, po.BC_WALLX])
def test_moments_after_steps_parity_bit_exact(bc):
    """the app's output sequence (app.f90:121-123) after 6 parity-mode steps:
    the fused entry point and the three routines both equal the oracle bit
    for bit"""
    from wumingpic_amd import MODE_PARITY
    p, up, uf, cc, np2 = _state(32, 24, 12, 17 + bc, bc)
    print(f"moments after steps, bc={bc}", flush=True)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.step(6)
        sim.moments()
        fused = sim.download_mom()
        sim.mom_accl()
        sim.mom_nvt()
        sim.bc_mom()
        split = sim.download_mom()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 6)
    o.mom_accl(gp, up, uf, cc)
    ref = o.mom_nvt(gp, np2)
    o.bc_mom(ref)
    assert np.array_equal(_bits(fused), _bits(ref))
    assert np.array_equal(_bits(split), _bits(ref))


def test_moments_fast_close_and_sum_rules():
    """fast mode on the seeded state: fused kick + deposit against the oracle,
    N summing to the particle count, and the split path agreeing"""
    from wumingpic_amd import MODE_FAST
    sim, p = _weibel(64, 32, 16, MODE_FAST)
    print("moments fast", flush=True)
    with sim:
        sim.step(3)
        up, np2, cc = sim.download_up(p.np)
        uf = sim.download_fields()
        sim.moments()
        fused = sim.download_mom()
        sim.mom_accl()
        sim.mom_nvt()
        sim.bc_mom()
        split = sim.download_mom()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, uf, cc)
    ref = o.mom_nvt(gp, np2)
    o.bc_mom(ref)
    _close(fused, ref)
    _close(split, ref)
    inner = fused[0, 1:p.nx + 1, 1:p.ny + 1, :]
    assert np.allclose(inner.sum(axis=(0, 1)), np2.sum(axis=0), rtol=1e-12)


@pytest.mark.parametrize("parity", [True, False])
def test_nvt_of_unsorted_gp(parity):
    """mom_calc__nvt accepts any gp: rows shuffled within themselves (the
    parity gather then scans whole rows, the fast path adds off-tile
    particles to HBM directly)"""
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    d, p = _golden()
    rng = np.random.default_rng(9)
    gp = np.array(d["gp"], order="F")
    np2 = np.asfortranarray(d["np2"])
    for s in range(p.nsp):
        for j in range(p.ny):
            n = np2[j, s]
            gp[:, :n, j, s] = gp[:, rng.permutation(n), j, s]
    print(f"nvt unsorted gp, parity={parity}", flush=True)
    with _pic(p, MODE_PARITY if parity else MODE_FAST) as sim:
        sim.upload_up(np.asfortranarray(d["up"]), np2, np.asfortranarray(d["cumcnt"]))
        sim.upload_gp(gp, np2)
        sim.mom_nvt()
        mom = sim.download_mom()
    ref = po.Oracle2D(p).mom_nvt(gp, np2)
    if parity:
        assert np.array_equal(_bits(mom), _bits(ref))
    else:
        _close(mom, ref)


def test_nvt_rejects_particle_outside_its_row():
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd._lib import WpicError
    d, p = _golden()
    gp = np.array(d["gp"], order="F")
    gp[1, 0, 3, 0] += 2.0  # two rows up from its own
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(np.asfortranarray(d["up"]), np.asfortranarray(d["np2"]), np.asfortranarray(d["cumcnt"]))
        sim.upload_gp(gp, np.asfortranarray(d["np2"]))
        with pytest.raises(WpicError):
            sim.mom_nvt()
