"""The shock boundaries on the device (2d/proj/shock/boundary_shock.f90, box fully expanded)
through the C-ABI: boundary_shock__injection and __particle_x bit-exact against the oracle, whole
steps of the shock app's loop order (2d/proj/shock/app.f90:111-116) bit-exact in parity mode and
within the tolerance written in the test in fast mode."""
import numpy as np
import pytest

from oracle import pyoracle as po
from shock_state import shock_state

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pic(p, mode):
    from wumingpic_amd.pic2d import Pic2D
    q = [p.q[k] for k in range(p.nsp)]
    r = [p.r[k] for k in range(p.nsp)]
    return Pic2D(p.nx, p.ny, q, r, nsp=p.nsp, delx=p.delx, delt=p.delt, c=p.c, gfac=p.gfac, bc=po.BC_SHOCK,
                 mode=mode, capacity=p.np * p.ny)


def _active_equal(a, b, np2):
    for s in range(np2.shape[1]):
        for j in range(np2.shape[0]):
            n = np2[j, s]
            if not np.array_equal(_bits(a[:, :n, j, s]), _bits(b[:, :n, j, s])):
                return False, (s, j)
    return True, None


@pytest.mark.parametrize("which", ["injection", "particle_x"])
def test_x_boundaries_match_oracle(which):
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = shock_state(40, 8, 8, 7, u0=-0.35)
    rng = np.random.default_rng(8)
    for s in range(2):
        for j in range(p.ny):
            n = np2[j, s]
            up[0, :n, j, s] = p.nxgs + rng.random(n) * (p.nx + 0.9)
    print(f"shock x boundary {which}", flush=True)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_gp(up, np2)
        if which == "injection":
            sim.bc_injection(p.u0)
        else:
            sim.bc_particle_x()
        gp, n2 = sim.download_gp(p.np)
    ref = up.copy(order="F")
    o = po.Oracle2D(p)
    if which == "injection":
        o.bc_injection(ref, np2)
    else:
        o.bc_particle_x(ref, np2)
    assert np.array_equal(n2, np2)
    ok, info = _active_equal(gp, ref, np2)
    assert ok, info


def test_shock_steps_parity_mode_bit_exact():
    from wumingpic_amd import MODE_PARITY
    p, up, uf, cc, np2 = shock_state(48, 16, 12, 9, u0=-0.35)
    print("shock steps parity", flush=True)
    with _pic(p, MODE_PARITY) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(p.u0)
        sim.step(12)
        u_dev, n2_dev, cc_dev = sim.download_up(p.np)
        uf_dev = sim.download_fields()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 12)
    assert np.array_equal(n2_dev, np2) and np.array_equal(cc_dev, cc)
    ok, info = _active_equal(u_dev, up, np2)
    assert ok, info
    assert np.array_equal(_bits(uf_dev), _bits(uf))


def test_shock_steps_fast_mode_close():
    """fast mode reorders the deposit's and CG's sums: fields within 1e-9 of their maximum"""
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = shock_state(48, 16, 12, 10, u0=-0.35)
    n0 = np2.sum(axis=0)
    print("shock steps fast", flush=True)
    with _pic(p, MODE_FAST) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_injection(p.u0)
        sim.step(12)
        uf_dev = sim.download_fields()
        counts = sim.counts()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 12)
    assert list(counts) == list(n0)
    assert np.abs(uf_dev - uf).max() <= 1e-9 * np.abs(uf).max()


def test_shock_rejected_in_the_mixed_form():
    from wumingpic_amd import MODE_MIXED, WpicError
    from wumingpic_amd.pic2d import Pic2D
    with pytest.raises(WpicError, match="FP64"):
        Pic2D(32, 16, [0.1, -0.1], [25.0, 1.0], bc=po.BC_SHOCK, mode=MODE_MIXED, capacity=1024)
