"""The oracle's particle__solv_vay (2d/common/particle.f90:182-315) against the reference's own
routine built from its sources (golden vectors written by tests/golden/gen_ref2d.py, and fresh
cases when the build is present), bit for bit, and against what the mathematics of the Vay kick
fixes: with B = 0 it is u + fac2 E; a particle at the relativistic E x B drift velocity keeps its
momentum (the property the scheme was made for, which the Boris kick lacks)."""
import os
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _params(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    return po.make_params(nx, ny, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def test_vay_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref2d_vay.npz"))
    p = _params(d)
    gp = np.zeros_like(d["up"], order="F")
    po.Oracle2D(p).particle_solv_vay(gp, np.asfortranarray(d["up"]), np.asfortranarray(d["uf"]),
                                     np.asfortranarray(d["cumcnt"]))
    assert np.array_equal(_bits(gp), _bits(d["gp"]))
    # the fixture's fields rotate strongly: the Vay and Boris kicks differ there
    gb = np.zeros_like(gp, order="F")
    po.Oracle2D(p).particle_solv(gb, np.asfortranarray(d["up"]), np.asfortranarray(d["uf"]),
                                 np.asfortranarray(d["cumcnt"]))
    n = int(d["np2"][0, 1])
    assert np.max(np.abs(gb[2:5, :n, 0, 1] - gp[2:5, :n, 0, 1])) > 1e-3


@pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built (needs the reference checkout)")
@pytest.mark.parametrize("seed", [4, 5])
def test_vay_fresh_cases_against_reference_build(seed):
    rng = np.random.default_rng(seed)
    nx, ny, n_ppc = 9, 7, 6
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(0, 12, size=(nx, ny)) for _ in range(2)]
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                     counts=counts)
    uf = S.random_fields(p, rng, scale=3.0 / abs(q[1]), b0=1.0 / abs(q[1]))
    with tempfile.TemporaryDirectory() as tmp:
        ref = po.run_ref2d(tmp, 4, p, up, uf=uf, cumcnt=cc)
    gp = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv_vay(gp, up, uf, cc)
    assert np.array_equal(_bits(gp), _bits(ref))


def _uniform_case(E, B, u, c=1.0):
    """one electron per cell of a 4x4 grid in uniform fields"""
    nx = ny = 4
    p = po.make_params(nx, ny, np_cap=4, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0), c=c)
    up = np.zeros((6, 4, ny, 2), order="F")
    cc = np.zeros((nx + 1, ny, 2), dtype=np.int32, order="F")
    for j in range(ny):
        cc[:, j, 1] = np.arange(nx + 1)
        for i in range(nx):
            up[0:2, i, j, 1] = [p.nxgs + i + 0.3 + 0.1 * j, p.nys + j + 0.7 - 0.1 * i]
            up[2:5, i, j, 1] = u
    uf = np.zeros((6, nx + 4, ny + 4), order="F")
    uf[0:3] = np.asarray(B)[:, None, None]
    uf[3:6] = np.asarray(E)[:, None, None]
    return p, up, uf, cc


def test_vay_without_magnetic_field_is_the_electric_kick():
    E, u = (0.3, -0.2, 0.05), np.array([0.4, 1.5, -0.7])
    p, up, uf, cc = _uniform_case(E, (0.0, 0.0, 0.0), u)
    gp = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv_vay(gp, up, uf, cc)
    fac2 = p.q[1] * p.delt / p.r[1]
    want = u + fac2 * np.array(E)
    got = gp[2:5, :4, :, 1].reshape(3, -1)
    assert np.allclose(got, want[:, None], rtol=4e-16 * 8, atol=0)
    gam = np.sqrt(1 + want @ want)
    assert np.allclose(gp[0, :4, :, 1], up[0, :4, :, 1] + want[0] * p.delt / gam, rtol=1e-15)


@pytest.mark.parametrize("vd", [0.3, 0.9, 0.99])
def test_vay_keeps_the_exb_drift(vd):
    """E = -v x B / c with v = vd c along y and B along z: the force vanishes, so the drift momentum
    u = gamma v stays; the Boris kick leaves it at relativistic vd (Vay 2008, section II)"""
    bz = 40.0  # fac1 * bz = 0.1 * 0.25 * 40 = 1: a strong field
    ex = vd * bz  # E + v x B / c = (ex + vy bz, 0, 0) = 0 with vy = -vd
    u = np.array([0.0, -vd / np.sqrt(1 - vd * vd), 0.0])
    p, up, uf, cc = _uniform_case((ex, 0.0, 0.0), (0.0, 0.0, bz), u)
    gp = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv_vay(gp, up, uf, cc)
    got = gp[2:5, :4, :, 1].reshape(3, -1)
    assert np.max(np.abs(got - u[:, None])) < 1e-12 * max(1.0, abs(u[1])) * 100
    gb = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv(gb, up, uf, cc)
    boris = gb[2:5, :4, :, 1].reshape(3, -1)
    if vd >= 0.9:
        assert np.max(np.abs(boris - u[:, None])) > 1e-3


def test_vay3d_golden_bit_exact():
    """3d/common/particle.f90:236-419 against the reference's own build"""
    d = np.load(os.path.join(GOLD, "ref3d_vay.npz"))
    nx, ny, nz = (int(d[f"n{a}ge"] - d[f"n{a}gs"] + 1) for a in "xyz")
    p = po.make_params3d(nx, ny, nz, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])
    gp = np.zeros_like(d["up"], order="F")
    po.Oracle3D(p).particle_solv_vay(gp, np.asfortranarray(d["up"]), np.asfortranarray(d["uf"]),
                                     np.asfortranarray(d["cumcnt"]))
    assert np.array_equal(_bits(gp), _bits(d["gp"]))
