"""The oracle against the reference's OWN particle__solv and sort__bucket
(golden vectors in tests/golden/, written by tests/golden/gen_ref2d.py from the
reference sources compiled with amdflang), bit for bit."""
import os
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _params(d):
    nx = int(d["nxge"] - d["nxgs"] + 1)
    ny = int(d["nye"] - d["nys"] + 1)
    return po.make_params(nx, ny, np_cap=int(d["np"]), delt=float(d["delt"]), q=d["q"], r=d["r"])


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_push_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref2d_push.npz"))
    p = _params(d)
    o = po.Oracle2D(p)
    gp = np.zeros_like(d["up"], order="F")
    o.particle_solv(gp, np.asfortranarray(d["up"]), np.asfortranarray(d["uf"]), np.asfortranarray(d["cumcnt"]))
    # whole array: active slots and the id slot copied over the full capacity
    assert np.array_equal(_bits(gp), _bits(d["gp"]))


def test_sort_golden_bit_exact():
    d = np.load(os.path.join(GOLD, "ref2d_sort.npz"))
    p = _params(d)
    o = po.Oracle2D(p)
    gp = np.zeros_like(d["up"], order="F")
    cc = np.zeros_like(d["cumcnt"], order="F")
    o.sort_bucket(gp, np.asfortranarray(d["up"]), cc, np.asfortranarray(d["np2"]))
    assert np.array_equal(cc, d["cumcnt"])
    np2 = d["np2"]
    for s in range(p.nsp):
        for j in range(p.ny):
            n = np2[j, s]
            assert np.array_equal(_bits(gp[:, :n, j, s]), _bits(d["gp"][:, :n, j, s]))


@pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built (needs the reference checkout)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_push_fresh_cases_against_reference_build(seed):
    """fresh random states through the reference's own particle__solv"""
    rng = np.random.default_rng(seed)
    nx, ny, n_ppc = 9, 7, 6
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    counts = [rng.integers(0, 12, size=(nx, ny)) for _ in range(2)]
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.3, 2.0), drift=(0.0, 0.5), colocate=False,
                                     counts=counts)
    uf = S.random_fields(p, rng, scale=1.0, b0=0.5)
    with tempfile.TemporaryDirectory() as tmp:
        ref = po.run_ref2d(tmp, 1, p, up, uf=uf, cumcnt=cc)
    gp = np.zeros_like(up, order="F")
    po.Oracle2D(p).particle_solv(gp, up, uf, cc)
    assert np.array_equal(_bits(gp), _bits(ref))


@pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built (needs the reference checkout)")
def test_sort_fresh_case_against_reference_build():
    rng = np.random.default_rng(9)
    nx, ny = 11, 6
    p = po.make_params(nx, ny, np_cap=90, delt=0.5, q=(0.1, -0.1), r=(25.0, 1.0))
    up = np.zeros((6, p.np, ny, 2), order="F")
    np2 = rng.integers(0, 90, size=(ny, 2)).astype(np.int32).copy(order="F")
    for s in range(2):
        for j in range(ny):
            n = np2[j, s]
            up[0, :n, j, s] = p.nxgs + nx * rng.random(n)
            up[1:5, :n, j, s] = rng.standard_normal((4, n))
            up[5, :n, j, s] = S.ids_as_f64(1 + np.arange(n) + 1000 * j + 100000 * s)
    with tempfile.TemporaryDirectory() as tmp:
        ref, ref_cc = po.run_ref2d(tmp, 2, p, up, np2=np2)
    gp = np.zeros_like(up, order="F")
    cc = np.zeros((nx + 1, ny, 2), dtype=np.int32, order="F")
    po.Oracle2D(p).sort_bucket(gp, up, cc, np2)
    assert np.array_equal(cc, ref_cc)
    for s in range(2):
        for j in range(ny):
            n = np2[j, s]
            assert np.array_equal(_bits(gp[:, :n, j, s]), _bits(ref[:, :n, j, s]))
