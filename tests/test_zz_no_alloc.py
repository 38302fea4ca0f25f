"""wpic_gauss2d_noalloc (the Gauss residual with its scratch in the library's
fixed device array, no allocation) against wpic_gauss2d on the same state:
two calls in a row on two contexts, the second grid smaller, each equal to
wpic_gauss2d's residual to 1e-12 of its maximum (the charge deposit adds with
atomics, so the two sums may differ in the last bits).  A grid beyond the
fixed scratch is rejected with a clear error and no GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(L):
    f = L.wpic_gauss2d_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")]
    return f


def _case(nx, ny, n_ppc, seed):
    """a state with ions and electrons drawn apart, so div E and rho are both
    O(1) and the residual is far above round-off"""
    from oracle import pyoracle as po
    import states2d as S
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd._lib import lib
    from wumingpic_amd.pic2d import Pic2D
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=False)
    uf = S.periodic_fill_fields(p, S.random_fields(p, rng, scale=0.05))
    L = lib()
    f = _noalloc(L)
    with Pic2D(nx, ny, list(q), list(r), delt=0.5, mode=MODE_FAST, capacity=2 * n_ppc * nx * ny) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        ref = sim.gauss()
        G = np.zeros((nx, ny), order="F")
        rc = f(sim._h, G)
        assert rc == 0, L.wpic_last_error()
        ref2 = sim.gauss()
    scale = np.abs(ref).max()
    assert scale > 1e-3
    assert np.abs(G - ref).max() <= 1e-12 * scale
    assert np.abs(ref2 - ref).max() <= 1e-12 * scale
    return G


def test_gauss_noalloc_matches_gauss_twice_second_smaller():
    print("wpic_gauss2d_noalloc: 96x64 then 40x32", flush=True)
    a = _case(96, 64, 8, 11)
    b = _case(40, 32, 8, 12)
    assert a.shape == (96, 64) and b.shape == (40, 32)


def test_gauss_noalloc_rejects_a_grid_beyond_its_scratch():
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd._lib import lib
    from wumingpic_amd.pic2d import Pic2D, weibel_charges
    L = lib()
    f = _noalloc(L)
    q, r = weibel_charges(1, 0.1, 25.0)
    with Pic2D(4096, 1024, q, r, delt=0.5, mode=MODE_FAST, capacity=16) as sim:
        G = np.zeros((1, 1), order="F")
        rc = f(sim._h, G)
        assert rc != 0
        L.wpic_last_error.restype = C.c_char_p
        assert b"exceeds the fixed scratch" in L.wpic_last_error()
