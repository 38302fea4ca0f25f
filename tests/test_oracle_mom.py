"""The oracle's moment routines: mom_calc__accl and mom_calc__nvt against the
reference's OWN mom_calc.f90 (2D and 3D golden vectors written by
tests/golden/gen_ref{2d,3d}.py from the reference sources compiled with
amdflang), bit for bit; the moment boundary folds (which need MPI in the
reference and are not built here) against a direct deposit with periodic or
mirrored node indices, and sum rules."""
import os
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S
import states3d as S3

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _p2(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    return po.make_params(nx, ny, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def _p3(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    nz = int(d["nze"] - d["nzs"] + 1)
    return po.make_params3d(nx, ny, nz, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def test_mom2d_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref2d_mom.npz"))
    p = _p2(d)
    o = po.Oracle2D(p)
    up = np.asfortranarray(d["up"])
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, np.asfortranarray(d["uf"]), np.asfortranarray(d["cumcnt"]))
    np2 = np.asfortranarray(d["np2"])
    for s in range(p.nsp):
        for j in range(p.ny):
            n = np2[j, s]
            assert np.array_equal(_bits(gp[:5, :n, j, s]), _bits(d["gp"][:5, :n, j, s]))
    mom = o.mom_nvt(np.asfortranarray(d["gp"]), np2)
    assert np.array_equal(_bits(mom), _bits(d["mom"]))


def test_mom3d_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref3d_mom.npz"))
    p = _p3(d)
    o = po.Oracle3D(p)
    up = np.asfortranarray(d["up"])
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, np.asfortranarray(d["uf"]), np.asfortranarray(d["cumcnt"]))
    np2 = np.asfortranarray(d["np2"])
    for s in range(p.nsp):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(gp[:6, :n, j, k, s]), _bits(d["gp"][:6, :n, j, k, s]))
    mom = o.mom_nvt(np.asfortranarray(d["gp"]), np2)
    assert np.array_equal(_bits(mom), _bits(d["mom"]))


@pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built (needs the reference checkout)")
@pytest.mark.parametrize("seed", [11, 12])
def test_mom2d_fresh_cases_against_reference_build(seed):
    rng = np.random.default_rng(seed)
    nx, ny, n_ppc = 9, 7, 6
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(0, 12, size=(nx, ny)) for _ in range(2)]
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                     counts=counts)
    uf = S.random_fields(p, rng, scale=0.5, b0=0.3)
    with tempfile.TemporaryDirectory() as tmp:
        gref, mref = po.run_ref2d(tmp, 3, p, up, np2=np2, uf=uf, cumcnt=cc)
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, uf, cc)
    for s in range(p.nsp):
        for j in range(p.ny):
            n = np2[j, s]
            assert np.array_equal(_bits(gp[:5, :n, j, s]), _bits(gref[:5, :n, j, s]))
    assert np.array_equal(_bits(o.mom_nvt(gp, np2)), _bits(mref))


@pytest.mark.skipif(not po.have_ref3d(), reason="oracle/_ref not built (needs the reference checkout)")
def test_mom3d_fresh_case_against_reference_build():
    rng = np.random.default_rng(13)
    nx, ny, nz, n_ppc = 6, 5, 4, 3
    q, r = S3.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=4 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(0, 7, size=(nx, ny, nz)) for _ in range(2)]
    up, cc, np2 = S3.sorted_particles3d(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                        counts=counts)
    uf = S3.random_fields3d(p, rng, scale=0.5, b0=0.3)
    with tempfile.TemporaryDirectory() as tmp:
        gref, mref = po.run_ref3d(tmp, 3, p, up, np2=np2, uf=uf, cumcnt=cc)
    o = po.Oracle3D(p)
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, uf, cc)
    for s in range(p.nsp):
        for k in range(p.nz):
            for j in range(p.ny):
                n = np2[j, k, s]
                assert np.array_equal(_bits(gp[:6, :n, j, k, s]), _bits(gref[:6, :n, j, k, s]))
    assert np.array_equal(_bits(o.mom_nvt(gp, np2)), _bits(mref))


def _moment_values(u, c):
    """the seven per-particle quantities N, u*gam, u^2*gam (mom_calc.f90:198-243)"""
    gam = 1.0 / np.sqrt(1.0 + (u ** 2).sum(axis=0) / (c * c))
    return np.concatenate([np.ones((1, u.shape[1])), u * gam, u * u * gam], axis=0)


def _direct2d(p, gp, np2, wall):
    """CIC deposit with the node index mapped straight onto the interior:
    periodic (x and y) or mirrored at the x walls (periodic y)."""
    nx, ny = p.nx, p.ny
    out = np.zeros((7, nx, ny, p.nsp))
    for s in range(p.nsp):
        for j in range(ny):
            n = np2[j, s]
            x, y, u = gp[0, :n, j, s], gp[1, :n, j, s], gp[2:5, :n, j, s]
            ih = np.floor(x - 0.5).astype(int)
            jh = np.floor(y - 0.5).astype(int)
            dx, dy = x - 0.5 - ih, y - 0.5 - jh
            m = _moment_values(u, p.c)
            for ox, wx in ((0, 1 - dx), (1, dx)):
                for oy, wy in ((0, 1 - dy), (1, dy)):
                    I = ih + ox - p.nxgs
                    if wall:
                        I = np.clip(I, 0, nx - 1)
                    else:
                        I = I % nx
                    J = (jh + oy - p.nys) % ny
                    for k in range(7):
                        np.add.at(out[k, :, :, s], (I, J), m[k] * wx * wy)
    return out


@pytest.mark.parametrize("wall", [False, True])
def test_bc_mom2d_equals_direct_folded_deposit(wall):
    rng = np.random.default_rng(31 + wall)
    nx, ny, n_ppc = 10, 8, 7
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r,
                       bc=po.BC_WALLX if wall else po.BC_PERIODIC)
    counts = [rng.integers(1, 12, size=(nx, ny)) for _ in range(2)]
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                     counts=counts)
    o = po.Oracle2D(p)
    mom = o.mom_nvt(up, np2)
    total = mom.sum(axis=(1, 2))
    o.bc_mom(mom)
    inner = mom[:, 1:nx + 1, 1:ny + 1, :]
    ref = _direct2d(p, up, np2, wall)
    assert np.allclose(inner, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # every ghost contribution lands on exactly one interior node
    assert np.allclose(inner.sum(axis=(1, 2)), total, rtol=1e-12)
    # N sums to the particle count: the four CIC weights sum to one
    assert np.allclose(inner[0].sum(axis=(0, 1)), np2.sum(axis=0), rtol=1e-13)


def test_bc_mom3d_equals_direct_folded_deposit():
    rng = np.random.default_rng(41)
    nx, ny, nz, n_ppc = 6, 5, 4, 3
    q, r = S3.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=4 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(1, 7, size=(nx, ny, nz)) for _ in range(2)]
    up, cc, np2 = S3.sorted_particles3d(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                        counts=counts)
    o = po.Oracle3D(p)
    mom = o.mom_nvt(up, np2)
    total = mom.sum(axis=(1, 2, 3))
    o.bc_mom(mom)
    inner = mom[:, 1:nx + 1, 1:ny + 1, 1:nz + 1, :]
    ref = np.zeros((7, nx, ny, nz, p.nsp))
    for s in range(p.nsp):
        for k in range(nz):
            for j in range(ny):
                n = np2[j, k, s]
                x = up[0:3, :n, j, k, s]
                h = np.floor(x - 0.5).astype(int)
                d = x - 0.5 - h
                m = _moment_values(up[3:6, :n, j, k, s], p.c)
                for corner in range(8):
                    o3 = [(corner >> a) & 1 for a in range(3)]
                    w = np.prod([d[a] if o3[a] else 1 - d[a] for a in range(3)], axis=0)
                    I = (h[0] + o3[0] - p.nxgs) % nx
                    J = (h[1] + o3[1] - p.nys) % ny
                    K = (h[2] + o3[2] - p.nzs) % nz
                    for l in range(7):
                        np.add.at(ref[l, :, :, :, s], (I, J, K), m[l] * w)
    assert np.allclose(inner, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(inner.sum(axis=(1, 2, 3)), total, rtol=1e-12)
    assert np.allclose(inner[0].sum(axis=(0, 1, 2)), np2.sum(axis=(0, 1)), rtol=1e-13)


def test_mom_accl_closed_forms():
    """zero fields leave gp = up; a uniform Bz rotates (ux, uy) by the Boris
    angle 2 atan(q B dt/4 / (r gamma)) of the half step and keeps |u|"""
    rng = np.random.default_rng(5)
    nx, ny, n_ppc = 6, 5, 4
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=(0.3, -0.2), r=(4.0, 1.0))
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.5, 1.0), drift=(0.0, 0.0))
    o = po.Oracle2D(p)
    uf = np.zeros((6, nx + 4, ny + 4), order="F")
    gp = np.zeros_like(up, order="F")
    o.mom_accl(gp, up, uf, cc)
    for s in range(2):
        for j in range(ny):
            n = np2[j, s]
            assert np.array_equal(gp[:5, :n, j, s], up[:5, :n, j, s])
    b = 0.7
    uf[2] = b
    o.mom_accl(gp, up, uf, cc)
    for s in range(2):
        fac1 = p.q[s] / p.r[s] * 0.5 * (p.delt * 0.5)
        for j in range(ny):
            n = np2[j, s]
            u, g = up[2:5, :n, j, s], gp[2:5, :n, j, s]
            gam = np.sqrt(p.c ** 2 + (u ** 2).sum(axis=0))
            t = fac1 * b / gam
            ang = 2 * np.arctan(t)
            rot_x = u[0] * np.cos(ang) + u[1] * np.sin(ang)
            rot_y = -u[0] * np.sin(ang) + u[1] * np.cos(ang)
            assert np.allclose(g[0], rot_x, rtol=1e-12, atol=1e-14)
            assert np.allclose(g[1], rot_y, rtol=1e-12, atol=1e-14)
            assert np.array_equal(g[2], u[2])
