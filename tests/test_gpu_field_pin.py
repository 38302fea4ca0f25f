"""The device field update against the closed form of one Fourier mode, the
device Gauss residual against the oracle's, and the reference's two abort
paths (field.f90:427-430 "stop at cgm after ite_max";
boundary_periodic.f90:231-234 "memory over (np2 > np)") forced from the host.

The mode model (tests/fourier_mode.py) is the time-centred Maxwell system on
the staggered grid, solved in numpy; it is independent of both the oracle and
the device code.  The CG runs at a relative tolerance of 1e-13 for the mode
checks; the abort test sets 1e-300, which rounding cannot meet, so every
kernel runs its normal course and the host stops after 100 iterations.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S
import states3d as S3
from fourier_mode import OFF2, OFF3, fields_of, mode_amplitudes, mode_step

pytestmark = pytest.mark.gpu


def _still_particles_2d(p):
    """one uncharged particle at rest per cell: the deposit adds exact zeros"""
    up, cc, np2 = S.sorted_particles(p, 1, np.random.default_rng(1), u_scale=(0.0, 0.0), drift=(0.0, 0.0))
    return up, cc, np2


@pytest.mark.parametrize("mode", ["fast", "parity"])
@pytest.mark.parametrize("m", [(1, 2), (3, -1)])
def test_device_field_update_fourier_mode_2d(mode, m):
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    from wumingpic_amd.pic2d import Pic2D
    nx, ny = 32, 24
    p = po.make_params(nx, ny, np_cap=4 * nx, delt=0.5, q=(0.0, 0.0), r=(25.0, 1.0))
    k = np.array([2 * np.pi * m[0] / nx, 2 * np.pi * m[1] / ny, 0.0])
    K = 2 * np.sin(k / 2)
    B, E, _ = mode_amplitudes(K, np.random.default_rng(40 + m[0]))
    J = np.zeros(3, dtype=complex)
    shape, lo = (nx + 4, ny + 4), (-2, -2)
    uf = fields_of(B, E, k[:2], OFF2, shape, lo)
    up, cc, np2 = _still_particles_2d(p)
    scale = np.abs(uf).max()
    with Pic2D(nx, ny, [0.0, 0.0], [25.0, 1.0], delt=0.5, mode=MODE_FAST if mode == "fast" else MODE_PARITY,
               capacity=nx * ny) as sim:
        sim.set_cg_tol(1e-13)
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        for step in range(3):
            sim.particle_solv()
            sim.field_fdtd_i()
            B, E = mode_step(K, B, E, J, p.c, p.delt, p.gfac)
            got = sim.download_fields()
            want = fields_of(B, E, k[:2], OFF2, shape, lo)
            err = np.abs(got - want).max()
            assert err <= 1e-11 * scale, (step, err)


@pytest.mark.parametrize("m", [(1, 2, 1), (2, -1, 3)])
def test_device_field_update_fourier_mode_3d(m):
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd.pic3d import Pic3D
    nx, ny, nz = 16, 12, 8
    p = po.make_params3d(nx, ny, nz, np_cap=4 * nx, delt=0.5, q=(0.0, 0.0), r=(25.0, 1.0))
    k = np.array([2 * np.pi * m[0] / nx, 2 * np.pi * m[1] / ny, 2 * np.pi * m[2] / nz])
    K = 2 * np.sin(k / 2)
    B, E, _ = mode_amplitudes(K, np.random.default_rng(50 + m[2]))
    J = np.zeros(3, dtype=complex)
    shape, lo = (nx + 4, ny + 4, nz + 4), (-2, -2, -2)
    uf = fields_of(B, E, k, OFF3, shape, lo)
    up, cc, np2 = S3.sorted_particles3d(p, 1, np.random.default_rng(2), u_scale=(0.0, 0.0), drift=(0.0, 0.0))
    scale = np.abs(uf).max()
    with Pic3D(nx, ny, nz, [0.0, 0.0], [25.0, 1.0], delt=0.5, mode=MODE_FAST, capacity=nx * ny * nz) as sim:
        sim.set_cg_tol(1e-13)
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        for step in range(2):
            sim.particle_solv()
            sim.field_fdtd_i()
            B, E = mode_step(K, B, E, J, p.c, p.delt, p.gfac)
            got = sim.download_fields()
            want = fields_of(B, E, k, OFF3, shape, lo)
            err = np.abs(got - want).max()
            assert err <= 1e-11 * scale, (step, err)


def _lumpy_state_2d(nx, ny, n_ppc, seed):
    rng = np.random.default_rng(seed)
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    # ions and electrons drawn apart: rho is O(q) per cell
    up, cc, np2 = S.sorted_particles(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=False)
    uf = S.periodic_fill_fields(p, S.random_fields(p, rng, scale=0.05))
    return p, up, uf, cc, np2


def test_device_gauss_equals_oracle_2d():
    """wpic_gauss2d == or2d_gauss on a state where both div E and rho are
    large (so a residual that is identically zero, or one built on the wrong
    shape or node, fails)"""
    from wumingpic_amd import MODE_FAST
    p, up, uf, cc, np2 = _lumpy_state_2d(32, 24, 8, 3)
    G_ref = po.Oracle2D(p).gauss(uf, up, np2)
    from wumingpic_amd.pic2d import Pic2D
    with Pic2D(p.nx, p.ny, [p.q[0], p.q[1]], [p.r[0], p.r[1]], delt=0.5, mode=MODE_FAST,
               capacity=p.np * p.ny) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        G = sim.gauss()
    G = np.asarray(G).reshape(G_ref.shape, order="F")
    assert np.abs(G_ref).max() > 0.1
    assert np.abs(G - G_ref).max() <= 1e-12 * np.abs(G_ref).max()


def test_device_gauss_held_over_fast_steps_2d():
    """20 fast-mode steps of a lumpy state: the device's G does not move
    (charge conservation to round-off) while div E changes by orders of
    magnitude more than G does"""
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd.pic2d import Pic2D
    p, up, uf, cc, np2 = _lumpy_state_2d(32, 24, 8, 4)
    with Pic2D(p.nx, p.ny, [p.q[0], p.q[1]], [p.r[0], p.r[1]], delt=0.5, mode=MODE_FAST,
               capacity=p.np * p.ny) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        G0 = np.asarray(sim.gauss()).copy()
        uf0 = sim.download_fields()
        sim.step(20)
        G1 = np.asarray(sim.gauss()).copy()
        uf1 = sim.download_fields()
    dive = lambda f: (f[3, 3:-1, 2:-2] - f[3, 2:-2, 2:-2]) + (f[4, 2:-2, 3:-1] - f[4, 2:-2, 2:-2])
    dG = np.abs(G1 - G0).max()
    assert dG <= 1e-11 * np.abs(G0).max()
    assert np.abs(dive(uf1) - dive(uf0)).max() > 1e4 * dG


def test_device_gauss_equals_oracle_3d():
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd.pic3d import Pic3D
    rng = np.random.default_rng(6)
    nx, ny, nz, n_ppc = 16, 12, 8, 4
    q, r = S.weibel_consts(n_ppc)
    p = po.make_params3d(nx, ny, nz, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    up, cc, np2 = S3.sorted_particles3d(p, n_ppc, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=False)
    uf = S3.periodic_fill_fields3d(p, S3.random_fields3d(p, rng, scale=0.05))
    G_ref = po.Oracle3D(p).gauss(uf, up, np2)
    with Pic3D(nx, ny, nz, [q[0], q[1]], [r[0], r[1]], delt=0.5, mode=MODE_FAST,
               capacity=p.np * ny * nz) as sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        G = np.asarray(sim.gauss()).reshape(G_ref.shape, order="F")
    assert np.abs(G_ref).max() > 0.1
    assert np.abs(G - G_ref).max() <= 1e-12 * np.abs(G_ref).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_cg_stop_after_ite_max(dim):
    """a tolerance rounding cannot meet: the call returns WPIC_ECG with the
    reference's text after 100 iterations, and the context still works"""
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd._lib import WpicError
    if dim == 2:
        from wumingpic_amd.pic2d import Pic2D
        p, up, uf, cc, np2 = _lumpy_state_2d(16, 16, 4, 8)
        sim = Pic2D(p.nx, p.ny, [p.q[0], p.q[1]], [p.r[0], p.r[1]], delt=0.5, mode=MODE_FAST,
                    capacity=p.np * p.ny)
    else:
        from wumingpic_amd.pic3d import Pic3D
        rng = np.random.default_rng(8)
        q, r = S.weibel_consts(2)
        p = po.make_params3d(8, 8, 8, np_cap=6 * 8, delt=0.5, q=q, r=r)
        up, cc, np2 = S3.sorted_particles3d(p, 2, rng, u_scale=(0.01, 0.1), drift=(0.0, 0.5), colocate=False)
        uf = S3.periodic_fill_fields3d(p, S3.random_fields3d(p, rng, scale=0.05))
        sim = Pic3D(8, 8, 8, [q[0], q[1]], [r[0], r[1]], delt=0.5, mode=MODE_FAST, capacity=p.np * 64)
    with sim:
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.set_cg_tol(1e-300)
        sim.particle_solv()
        with pytest.raises(WpicError) as ei:
            sim.field_fdtd_i()
        assert ei.value.code == -3  # WPIC_ECG
        assert "stop at cgm after ite_max" in str(ei.value)
        # the kernels ended normally: the context keeps working
        sim.set_cg_tol(1e-6)
        sim.upload_up(up, np2, cc)
        sim.upload_fields(uf)
        sim.particle_solv()
        sim.field_fdtd_i()


def test_memory_over_paths():
    """np2 > np: a download into rows narrower than a row's count, an upload
    beyond the device capacity and an injection beyond it all return
    WPIC_EMEM with the reference's text"""
    from wumingpic_amd import MODE_FAST
    from wumingpic_amd._lib import WpicError
    from wumingpic_amd.pic2d import Pic2D
    p, up, uf, cc, np2 = _lumpy_state_2d(16, 16, 4, 9)
    with Pic2D(p.nx, p.ny, [p.q[0], p.q[1]], [p.r[0], p.r[1]], delt=0.5, mode=MODE_FAST,
               capacity=4 * 16 * 16) as sim:
        sim.upload_up(up, np2, cc)
        with pytest.raises(WpicError, match=r"memory over \(np2 > np\)"):
            sim.download_up(int(np2.max()) - 1)
        big = np2.copy(order="F")
        big[0, 0] += 1  # one particle more than the species capacity
        up2 = np.zeros((6, p.np, p.ny, 2), order="F")
        up2[:, :, :, :] = up
        up2[:, int(np2[0, 0]), 0, 0] = up[:, int(np2[0, 0]) - 1, 0, 0]  # a second copy in the row's last cell
        cc2 = cc.copy(order="F")
        cc2[-1, 0, 0] += 1  # keep cumcnt(nxe+1) = np2: the capacity check is the one that fires
        with pytest.raises(WpicError, match=r"memory over \(np2 > np\)"):
            sim.upload_up(up2, big, cc2)
