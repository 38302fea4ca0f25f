"""The 3D oracle (oracle/oracle3d.c) pinned three ways:

  push, sort       bit for bit against the reference's OWN 3d/common/particle.f90
                   and sort.f90 (golden vectors from tests/golden/gen_ref3d.py;
                   fresh cases when oracle/_ref is built);
  Esirkepov        discrete continuity rho1 - rho0 + dt div J = 0 in every cell
                   (a dropped term, a wrong pj permutation on the scatter or a
                   wrong dstmp factor breaks it);
  CG               an FFT solve of the periodic 7-point operator f4 - 2 sum cos;
  whole step       Gauss's law stays at round-off (exercises the curl and the
                   ghost fills of dfield and curre), particles, ids and column
                   membership are conserved, columns stay sorted by x cell.
The y-z exchange's in-column order after hole filling
(boundary_periodic__particle_yz) has no reference build here (the module needs
MPI): it is parity unpinned; its multiset and membership are checked.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
import states3d as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _params(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    nz = int(d["nze"] - d["nzs"] + 1)
    return po.make_params3d(nx, ny, nz, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _active(up, np2):
    ny, nz, nsp = np2.shape
    return np.concatenate([up[:, :np2[j, k, s], j, k, s] for s in range(nsp) for k in range(nz) for j in range(ny)],
                          axis=1)


def test_push3d_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref3d_push.npz"))
    p = _params(d)
    gp = np.zeros_like(d["up"], order="F")
    po.Oracle3D(p).particle_solv(gp, np.asfortranarray(d["up"]), np.asfortranarray(d["uf"]),
                                 np.asfortranarray(d["cumcnt"]))
    assert np.array_equal(_bits(gp), _bits(d["gp"]))


def test_sort3d_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref3d_sort.npz"))
    p = _params(d)
    gp = np.zeros_like(d["up"], order="F")
    cc = np.zeros_like(d["cumcnt"], order="F")
    po.Oracle3D(p).sort_bucket(gp, np.asfortranarray(d["up"]), cc, np.asfortranarray(d["np2"]))
    assert np.array_equal(cc, d["cumcnt"])
    np2 = d["np2"]
    for s in range(p.nsp):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(gp[:, :n, j, k, s]), _bits(d["gp"][:, :n, j, k, s]))


@pytest.mark.skipif(not po.have_ref3d(), reason="oracle/_ref not built (needs the reference checkout)")
@pytest.mark.parametrize("seed", [1, 2])
def test_push3d_fresh_cases_against_reference_build(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz, n_ppc = 7, 5, 4, 3
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(0, 7, size=(nx, ny, nz)) for _ in range(2)]
    up, cc, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                       counts=counts)
    uf = S.random_fields3d(p, rng, scale=1.0, b0=0.5)
    with tempfile.TemporaryDirectory() as tmp:
        ref = po.run_ref3d(tmp, 1, p, up, uf=uf, cumcnt=cc)
    gp = np.zeros_like(up, order="F")
    po.Oracle3D(p).particle_solv(gp, up, uf, cc)
    assert np.array_equal(_bits(gp), _bits(ref))


def test_boris3d_uniform_b_gyration():
    """constant fields interpolate exactly (the 27-point quadratic weights sum
    to one): |u| is kept and u rotates by 2 atan(fac1 |B| / gamma)"""
    rng = np.random.default_rng(5)
    p = po.make_params3d(6, 5, 4, np_cap=40, delt=0.5, q=(0.3, -0.3), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles3d(p, 3, rng, u_scale=(0.2, 0.5), drift=(0.0, 0.0), colocate=False)
    uf = np.zeros((6, p.nx + 4, p.ny + 4, p.nz + 4), order="F")
    uf[2] = 0.8
    gp = np.zeros_like(up, order="F")
    po.Oracle3D(p).particle_solv(gp, up, uf, cc)
    a0, a1 = _active(up, np2), _active(gp, np2)
    n0 = np.linalg.norm(a0[3:6], axis=0)
    n1 = np.linalg.norm(a1[3:6], axis=0)
    np.testing.assert_allclose(n1, n0, rtol=1e-14)
    np.testing.assert_array_equal(a1[5], a0[5])  # B along z leaves u_z alone
    qm = np.where(np.arange(a0.shape[1]) < np2[..., 0].sum(), 0.3 / 25.0, -0.3)
    gam = np.sqrt(1.0 + n0 ** 2)
    ang = np.arctan2(a1[4], a1[3]) - np.arctan2(a0[4], a0[3])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    np.testing.assert_allclose(ang, -2 * np.arctan(qm * 0.25 * 0.8 / gam), rtol=1e-10, atol=1e-14)
    # the move uses the new u: x1 = x0 + u1 dt / gamma1
    np.testing.assert_allclose(a1[0], a0[0] + a1[3] * 0.5 / np.sqrt(1 + n1 ** 2), rtol=0, atol=1e-13)


def _rho_centres3d(p, xyz, q):
    nx, ny, nz = p.nx, p.ny, p.nz
    rho = np.zeros((nx, ny, nz))
    c0 = np.floor(xyz).astype(int)
    d = xyz - 0.5 - c0
    sh = [[0.5 * (0.5 - d[a]) ** 2, 0.75 - d[a] ** 2, 0.5 * (0.5 + d[a]) ** 2] for a in range(3)]
    lo = (p.nxgs, p.nys, p.nzs)
    n = (nx, ny, nz)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                idx = tuple((c0[ax] - 1 + o - lo[ax]) % n[ax] for ax, o in zip(range(3), (a, b, c)))
                np.add.at(rho, idx, q * sh[0][a] * sh[1][b] * sh[2][c])
    return rho


def test_deposit3d_discrete_continuity():
    rng = np.random.default_rng(4)
    p = po.make_params3d(7, 6, 5, np_cap=40, delt=0.5, q=(0.37, -0.21), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles3d(p, 3, rng, u_scale=(0.1, 0.1), drift=(0.0, 0.0), colocate=False)
    gp = up.copy(order="F")
    for s in range(2):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                gp[0:3, :n, j, k, s] += rng.uniform(-0.49, 0.49, (3, n))
    o = po.Oracle3D(p)
    uj = np.zeros((3, p.nx + 4, p.ny + 4, p.nz + 4), order="F")
    o.ele_cur(uj, up, gp, cc)
    o.bc_curre(uj)
    drho = np.zeros((p.nx, p.ny, p.nz))
    for s in range(2):
        drho += _rho_centres3d(p, _active_species(gp, np2, s)[:3], p.q[s]) - \
                _rho_centres3d(p, _active_species(up, np2, s)[:3], p.q[s])
    nx, ny, nz = p.nx, p.ny, p.nz
    I, J, K = slice(2, nx + 2), slice(2, ny + 2), slice(2, nz + 2)
    divj = (uj[0, 3:nx + 3, J, K] - uj[0, I, J, K]) + (uj[1, I, 3:ny + 3, K] - uj[1, I, J, K]) + \
           (uj[2, I, J, 3:nz + 3] - uj[2, I, J, K])
    resid = drho + p.delt * divj
    assert np.abs(drho).max() > 1e-3
    assert np.abs(resid).max() < 1e-15 * 64


def _active_species(up, np2, s):
    ny, nz, _ = np2.shape
    return np.concatenate([up[:, :np2[j, k, s], j, k, s] for k in range(nz) for j in range(ny)], axis=1)


def test_deposit3d_particle_at_rest_gives_no_current():
    rng = np.random.default_rng(6)
    p = po.make_params3d(6, 5, 4, np_cap=30, delt=0.5, q=(0.5, -0.5), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles3d(p, 2, rng, u_scale=(0.1, 0.1), drift=(0.0, 0.0), colocate=False)
    uj = np.ones((3, p.nx + 4, p.ny + 4, p.nz + 4), order="F")
    po.Oracle3D(p).ele_cur(uj, up, up.copy(order="F"), cc)
    assert np.abs(uj).max() == 0.0


def test_cg3d_matches_fft_solve_periodic():
    rng = np.random.default_rng(12)
    p = po.make_params3d(10, 8, 6, np_cap=10, delt=0.5, q=(1.0, -1.0), r=(25.0, 1.0))
    o = po.Oracle3D(p)
    f = o.field_consts()
    f4, f5 = f[3], f[4]
    assert f4 == 6.0 + f5
    o.gkl[:] = rng.standard_normal(o.gkl.shape)
    o.df[:] = 0.0
    iters = o.cgm()
    assert min(iters) >= 1
    k = [2 * np.pi * np.fft.fftfreq(n) for n in (p.nx, p.ny, p.nz)]
    sym = f4 - 2 * np.cos(k[0])[:, None, None] - 2 * np.cos(k[1])[None, :, None] - 2 * np.cos(k[2])[None, None, :]
    for l in range(3):
        b = f5 * o.gkl[l]
        phi = np.real(np.fft.ifftn(np.fft.fftn(b) / sym))
        got = o.df[l, 2:p.nx + 2, 2:p.ny + 2, 2:p.nz + 2]
        assert np.abs(got - phi).max() <= 1e-5 * np.abs(phi).max()


def test_bc_particle_yz_conserves_and_places():
    rng = np.random.default_rng(21)
    p = po.make_params3d(6, 5, 4, np_cap=60, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    up, cc, np2 = S.sorted_particles3d(p, 4, rng, u_scale=(0.1, 0.1), drift=(0.0, 0.0), colocate=False)
    # move a third of the particles by up to one cell in y and z (wrapping
    # across the periodic faces), and x across the periodic x faces
    for s in range(2):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                m = rng.random(n) < 0.35
                up[1, :n, j, k, s] += m * rng.uniform(-0.9, 0.9, n)
                up[2, :n, j, k, s] += m * rng.uniform(-0.9, 0.9, n)
    ids0 = np.sort(_active(up, np2)[6].view(np.int64))
    o = po.Oracle3D(p)
    np2b = np2.copy(order="F")
    o.bc_particle_x(up, np2b)
    o.bc_particle_yz(up, np2b)
    a = _active(up, np2b)
    assert np.array_equal(np.sort(a[6].view(np.int64)), ids0)
    ly, lz = p.ny * p.delx, p.nz * p.delx
    for s in range(2):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2b[j, k, s]
                assert (np.floor(up[1, :n, j, k, s]).astype(int) == p.nys + j).all()
                assert (np.floor(up[2, :n, j, k, s]).astype(int) == p.nzs + k).all()
    assert (a[1] >= p.nys).all() and (a[1] < p.nys + ly).all()
    assert (a[2] >= p.nzs).all() and (a[2] < p.nzs + lz).all()


def _weibel3d_state(nx, ny, nz, n_ppc, seed=20261202):
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S.sorted_particles3d(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=True)
    uf = np.zeros((6, nx + 4, ny + 4, nz + 4), order="F")
    return p, up, uf, cc, np2


def test_full_step3d_conservation_and_gauss_law():
    p, up, uf, cc, np2 = _weibel3d_state(8, 8, 6, 4)
    o = po.Oracle3D(p)
    ids0 = np.sort(_active(up, np2)[6].view(np.int64))
    G0 = o.gauss(uf, up, np2)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 8)
    G1 = o.gauss(uf, up, np2)
    scale = 4 * np.pi * abs(p.q[1]) * 4
    assert np.abs(G0).max() < 1e-13 * scale
    assert np.abs(G1).max() < 1e-12 * scale
    assert np.abs(uf).max() > 1e-8
    assert np.array_equal(np.sort(_active(up, np2)[6].view(np.int64)), ids0)
    assert np2.sum() == 2 * 4 * 8 * 8 * 6
    for s in range(2):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                x, y, z = up[0, :n, j, k, s], up[1, :n, j, k, s], up[2, :n, j, k, s]
                assert (np.floor(y).astype(int) == p.nys + j).all()
                assert (np.floor(z).astype(int) == p.nzs + k).all()
                cells = np.floor(x).astype(int)
                assert (np.diff(cells) >= 0).all()
                assert np.array_equal(np.bincount(cells - p.nxgs, minlength=p.nx), np.diff(cc[:, j, k, s]))
