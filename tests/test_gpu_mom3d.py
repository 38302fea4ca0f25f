"""GPU parity of the 3D moments (mom_calc__accl, mom_calc__nvt,
boundary_periodic__mom) through the C-ABI, against the reference's own golden
vectors and the 3D oracle.

Bar: mom_calc__accl bit-exact in every mode; mom_calc__nvt and the folds
bit-exact in parity mode; fast mode within 1e-13 of each moment's largest
magnitude.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
import states3d as S

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FAST_TOL = 1e-13


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pic(p, mode):
    from wumingpic_amd.pic3d import Pic3D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    return Pic3D(p.nx, p.ny, p.nz, q, r, nsp=p.nsp, delt=p.delt, mode=mode, capacity=p.np * p.ny * p.nz)


def _close(a, b):
    for k in range(7):
        scale = max(np.abs(b[k]).max(), 1e-300)
        err = np.abs(a[k] - b[k]).max() / scale
        assert err <= FAST_TOL, (k, err)


def _golden():
    d = np.load(os.path.join(GOLD, "ref3d_mom.npz"))
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    nz = int(d["nze"] - d["nzs"] + 1)
    p = po.make_params3d(nx, ny, nz, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])
    return d, p


@pytest.mark.parametrize("parity", [True, False])
def test_accl3d_and_nvt3d_match_reference_golden(parity):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    d, p = _golden()
    np2 = d["np2"]
    print(f"3d accl+nvt golden, parity={parity}", flush=True)
    with _pic(p, MODE_PARITY if parity else MODE_FAST) as sim:
        sim.upload_up(d["up"], np2, d["cumcnt"])
        sim.upload_fields(d["uf"])
        sim.mom_accl()
        gp, n2 = sim.download_gp(p.np)
        sim.mom_nvt()
        mom = sim.download_mom()
    assert np.array_equal(n2, np2)
    for s in range(p.nsp):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(gp[:6, :n, j, k, s]), _bits(d["gp"][:6, :n, j, k, s]))
                assert np.array_equal(_bits(gp[6, :n, j, k, s]), _bits(d["up"][6, :n, j, k, s]))
    if parity:
        assert np.array_equal(_bits(mom), _bits(d["mom"]))
    else:
        _close(mom, d["mom"])


def test_bc_mom3d_matches_oracle():
    from wumingpic_amd import MODE_PARITY
    rng = np.random.default_rng(4)
    p = po.make_params3d(9, 6, 5, np_cap=64, q=(0.1, -0.1), r=(25.0, 1.0))
    mom = np.asfortranarray(rng.standard_normal((7, p.nx + 2, p.ny + 2, p.nz + 2, p.nsp)))
    print("3d bc_mom", flush=True)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_mom(mom)
        sim.bc_mom()
        dev = sim.download_mom()
    po.Oracle3D(p).bc_mom(mom)
    assert np.array_equal(_bits(dev), _bits(mom))


def _weibel3d(nx, ny, nz, n_ppc, mode):
    from wumingpic_amd.pic3d import weibel_beams_3d
    sim = weibel_beams_3d(nx, ny, nz, n_ppc, mode=mode, capacity=2 * n_ppc * nx * ny * nz)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    return sim, p


def test_moments3d_after_steps_parity_bit_exact():
    """the output sequence (3d/proj/weibel/app.f90:122-124) after 4 parity
    steps, fused and as three routines, bit for bit against the oracle"""
    from wumingpic_amd import MODE_PARITY
    sim, p = _weibel3d(12, 8, 6, 4, MODE_PARITY)
    print("3d moments after steps", flush=True)
    with sim:
        up, np2, cc = sim.download_up(p.np)
        uf = sim.download_fields()
        sim.step(4)
        sim.moments()
        fused = sim.download_mom()
        sim.mom_accl()
        sim.mom_nvt()
        sim.bc_mom()
        split = sim.download_mom()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 4)
    o.mom_accl(gp, up, uf, cc)
    ref = o.mom_nvt(gp, np2)
    o.bc_mom(ref)
    assert np.array_equal(_bits(fused), _bits(ref))
    assert np.array_equal(_bits(split), _bits(ref))


def test_moments3d_fast_close_and_sum_rules():
    from wumingpic_amd import MODE_FAST
    sim, p = _weibel3d(24, 16, 12, 4, MODE_FAST)
    print("3d moments fast", flush=True)
    with sim:
        sim.step(2)
        up, np2, cc = sim.download_up(p.np)
        uf = sim.download_fields()
        sim.moments()
        fused = sim.download_mom()
        sim.mom_accl()
        sim.mom_nvt()
        sim.bc_mom()
        split = sim.download_mom()
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, uf, cc)
    ref = o.mom_nvt(gp, np2)
    o.bc_mom(ref)
    _close(fused, ref)
    _close(split, ref)
    inner = fused[0, 1:p.nx + 1, 1:p.ny + 1, 1:p.nz + 1, :]
    assert np.allclose(inner.sum(axis=(0, 1, 2)), np2.sum(axis=(0, 1)), rtol=1e-12)


@pytest.mark.parametrize("parity", [True, False])
def test_nvt3d_of_unsorted_gp(parity):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    d, p = _golden()
    rng = np.random.default_rng(10)
    gp = np.array(d["gp"], order="F")
    np2 = d["np2"]
    for s in range(p.nsp):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                gp[:, :n, j, k, s] = gp[:, rng.permutation(n), j, k, s]
    print(f"3d nvt unsorted gp, parity={parity}", flush=True)
    with _pic(p, MODE_PARITY if parity else MODE_FAST) as sim:
        sim.upload_up(d["up"], np2, d["cumcnt"])
        sim.upload_gp(gp, np2)
        sim.mom_nvt()
        mom = sim.download_mom()
    ref = po.Oracle3D(p).mom_nvt(gp, np2)
    if parity:
        assert np.array_equal(_bits(mom), _bits(ref))
    else:
        _close(mom, ref)
