"""The 3D mixed-precision form (WPIC_MODE_MIXED: int32 x cell + FP32 offsets
and momenta, FP32 push/deposit with FP32 LDS accumulation, FP64 global
current sum and FP64 field solve) against the 3D FP64 path from the same
seeded state.

Tolerances written here, as for the 2D form: positions to 2e-6 of a cell and
momenta to 2e-6 relative after one step; fields to 1e-4 of their maximum over
the first steps; Gauss's law to FP32 round-off of the deposit.
"""
import numpy as np
import pytest

import states2d as S

pytestmark = pytest.mark.gpu


def _pair(nx, ny, nz, n_ppc, seed=20261202):
    from wumingpic_amd import MODE_FAST, MODE_MIXED
    from wumingpic_amd.pic3d import Pic3D, weibel_charges
    q, r = weibel_charges(n_ppc, 0.1, 25.0)
    sims = []
    for mode in (MODE_FAST, MODE_MIXED):
        sim = Pic3D(nx, ny, nz, q, r, delt=0.5, mode=mode, capacity=2 * n_ppc * nx * ny * nz)
        sim.ic_beams(seed, n_ppc, 0.5, 0.1, 0.01)
        sims.append(sim)
    return sims


def _by_id(up, np2, j, k, s):
    n = np2[j, k, s]
    p = up[:, :n, j, k, s]
    return p[:, np.argsort(p[6].view(np.int64))]


def test_mixed3d_initial_state_is_the_fp64_state_rounded():
    f64, mix = _pair(12, 8, 6, 8)
    cap = 3 * 8 * 12
    with f64, mix:
        a, n2a, cca = f64.download_up(cap)
        b, n2b, ccb = mix.download_up(cap)
    assert np.array_equal(n2a, n2b) and np.array_equal(cca, ccb)
    for s in range(2):
        for k in range(6):
            for j in range(8):
                n = n2a[j, k, s]
                assert np.array_equal(np.floor(a[0:3, :n, j, k, s]), np.floor(b[0:3, :n, j, k, s]))
                assert np.abs(a[0:3, :n, j, k, s] - b[0:3, :n, j, k, s]).max() <= 1e-7
                u = a[3:6, :n, j, k, s]
                assert np.abs(u - b[3:6, :n, j, k, s]).max() <= 1e-7 * max(1.0, np.abs(u).max())
                assert np.array_equal(a[6, :n, j, k, s].view(np.int64), b[6, :n, j, k, s].view(np.int64))


def test_mixed3d_one_step_matches_fp64():
    f64, mix = _pair(12, 8, 6, 8)
    cap = 3 * 8 * 12
    with f64, mix:
        f64.step(1)
        mix.step(1)
        a, n2a, _ = f64.download_up(cap)
        b, n2b, _ = mix.download_up(cap)
        ufa, ufb = f64.download_fields(), mix.download_fields()
        assert list(mix.counts()) == [8 * 12 * 8 * 6] * 2
    # the same particles in the same columns; a particle within FP32 rounding
    # of a y or z face could sit one column over, which these seeds avoid
    assert np.array_equal(n2a, n2b)
    for s in range(2):
        for k in range(6):
            for j in range(8):
                pa, pb = _by_id(a, n2a, j, k, s), _by_id(b, n2b, j, k, s)
                assert np.array_equal(pa[6].view(np.int64), pb[6].view(np.int64))
                assert np.abs(pa[0:3] - pb[0:3]).max() <= 2e-6
                assert np.abs(pa[3:6] - pb[3:6]).max() <= 2e-6 * max(1.0, np.abs(pa[3:6]).max())
    scale = np.abs(ufa).max()
    assert scale > 0
    assert np.abs(ufa - ufb).max() <= 1e-4 * scale


def test_mixed3d_steps_conserve_and_track_fp64():
    f64, mix = _pair(12, 8, 6, 8)
    with f64, mix:
        G0 = mix.gauss()
        f64.step(12)
        mix.step(12)
        ufa, ufb = f64.download_fields(), mix.download_fields()
        G1 = mix.gauss()
        counts = mix.counts()
        ea, eb = f64.energy(), mix.energy()
    assert list(counts) == [8 * 12 * 8 * 6] * 2
    q, r = S.weibel_consts(8)
    rho_scale = 4 * np.pi * abs(q[1]) * 8
    assert np.abs(G1).max() <= 1e-5 * rho_scale + np.abs(G0).max() * 4
    assert np.abs(ufa - ufb).max() <= 1e-4 * np.abs(ufa).max()
    assert abs(ea[0] - eb[0]) <= 1e-5 * ea[0]


def test_mixed3d_rejects_single_routine_entry_points():
    from wumingpic_amd import MODE_FAST, WpicError
    f64, mix = _pair(8, 4, 4, 2)
    with f64, mix:
        with pytest.raises(WpicError, match="mixed-precision"):
            mix.particle_solv()
        with pytest.raises(WpicError, match="mixed-precision"):
            mix.sort_bucket()
        with pytest.raises(WpicError, match="fixed when the context is created"):
            mix.set_mode(MODE_FAST)
