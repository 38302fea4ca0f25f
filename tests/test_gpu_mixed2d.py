"""The mixed-precision form (WPIC_MODE_MIXED: int32 x cell + FP32 offsets and
momenta, FP32 push/deposit with FP32 LDS accumulation, FP64 global current
sum and FP64 field solve) against the FP64 path from the same seeded state.

Tolerances written here: positions to 2e-6 of a cell and momenta to 2e-6
relative after one step (FP32 rounding of a handful of operations); fields to
1e-4 of their maximum over the first steps, where round-off has not yet been
amplified by the instability; Gauss's law to FP32 round-off of the deposit.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
import states2d as S

pytestmark = pytest.mark.gpu


def _pair(nx, ny, n_ppc, bc=0, seed=20261202):
    from wumingpic_amd import MODE_FAST, MODE_MIXED
    from wumingpic_amd.pic2d import Pic2D, weibel_charges
    q, r = weibel_charges(n_ppc, 0.1, 25.0)
    sims = []
    for mode in (MODE_FAST, MODE_MIXED):
        sim = Pic2D(nx, ny, q, r, delt=0.5, mode=mode, bc=bc, capacity=2 * n_ppc * nx * ny)
        sim.ic_beams(seed, n_ppc, 0.5, 0.1, 0.01, 0.0)
        sims.append(sim)
    return sims


def _sorted_state(sim, np_cap):
    up, np2, cc = sim.download_up(np_cap)
    return up, np2, cc


def test_mixed_initial_state_is_the_fp64_state_rounded():
    f64, mix = _pair(32, 32, 16)
    with f64, mix:
        a, n2a, cca = _sorted_state(f64, 3 * 16 * 32)
        b, n2b, ccb = _sorted_state(mix, 3 * 16 * 32)
    assert np.array_equal(n2a, n2b) and np.array_equal(cca, ccb)
    for s in range(2):
        for j in range(32):
            n = n2a[j, s]
            assert np.array_equal(np.floor(a[0, :n, j, s]), np.floor(b[0, :n, j, s]))
            assert np.abs(a[0:2, :n, j, s] - b[0:2, :n, j, s]).max() <= 1e-7
            assert np.abs(a[2:5, :n, j, s] - b[2:5, :n, j, s]).max() <= 1e-7 * max(1.0, np.abs(a[2:5, :n, j, s]).max())
            assert np.array_equal(a[5, :n, j, s].view(np.int64), b[5, :n, j, s].view(np.int64))


@pytest.mark.parametrize("bc", [po.BC_PERIODIC, po.BC_WALLX])
def test_mixed_one_step_matches_fp64(bc):
    f64, mix = _pair(32, 24, 16, bc=bc)
    with f64, mix:
        f64.step(1)
        mix.step(1)
        a, n2a, cca = _sorted_state(f64, 3 * 16 * 32)
        b, n2b, ccb = _sorted_state(mix, 3 * 16 * 32)
        ufa, ufb = f64.download_fields(), mix.download_fields()
        assert list(mix.counts()) == [16 * 32 * 24] * 2
    assert np.array_equal(n2a, n2b)
    # the same particles in the same rows, in cell order; a particle within
    # FP32 rounding of a face may sit one cell over
    ids_a = np.concatenate([np.sort(a[5, :n2a[j, s], j, s].view(np.int64)) for s in range(2) for j in range(24)])
    ids_b = np.concatenate([np.sort(b[5, :n2b[j, s], j, s].view(np.int64)) for s in range(2) for j in range(24)])
    assert np.array_equal(ids_a, ids_b)
    for s in range(2):
        for j in range(24):
            n = n2a[j, s]
            ia = np.argsort(a[5, :n, j, s].view(np.int64))
            ib = np.argsort(b[5, :n, j, s].view(np.int64))
            pa, pb = a[:, :n, j, s][:, ia], b[:, :n, j, s][:, ib]
            assert np.abs(pa[0:2] - pb[0:2]).max() <= 2e-6
            assert np.abs(pa[2:5] - pb[2:5]).max() <= 2e-6 * max(1.0, np.abs(pa[2:5]).max())
    scale = np.abs(ufa).max()
    assert scale > 0
    assert np.abs(ufa - ufb).max() <= 1e-4 * scale


def test_mixed_steps_conserve_and_track_fp64():
    f64, mix = _pair(32, 32, 16)
    with f64, mix:
        G0 = mix.gauss()
        f64.step(20)
        mix.step(20)
        ufa, ufb = f64.download_fields(), mix.download_fields()
        G1 = mix.gauss()
        counts = mix.counts()
        ea, eb = f64.energy(), mix.energy()
    assert list(counts) == [16 * 32 * 32] * 2
    q, r = S.weibel_consts(16)
    rho_scale = 4 * np.pi * abs(q[1]) * 16
    # FP32 deposit: continuity holds to FP32 round-off of the weights
    assert np.abs(G1).max() <= 1e-5 * rho_scale + np.abs(G0).max() * 4
    assert np.abs(ufa - ufb).max() <= 1e-4 * np.abs(ufa).max()
    assert abs(ea[0] - eb[0]) <= 1e-5 * ea[0]


def test_mixed_rejects_single_routine_entry_points():
    from wumingpic_amd import MODE_FAST, WpicError
    f64, mix = _pair(16, 16, 4)
    with f64, mix:
        with pytest.raises(WpicError, match="mixed-precision"):
            mix.particle_solv()
        with pytest.raises(WpicError, match="fixed when the context is created"):
            mix.set_mode(MODE_FAST)


MOMENTS = ("N", "Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz")


def _moment_errors(mom, ref):
    """max |mom - ref| / max |ref| per moment and species"""
    err = {}
    for s in range(ref.shape[3]):
        for k, name in enumerate(MOMENTS):
            scale = np.abs(ref[k, :, :, s]).max()
            err[f"{name}[{s}]"] = float(np.abs(mom[k, :, :, s] - ref[k, :, :, s]).max() / scale)
    return err


@pytest.mark.parametrize("nsteps,tol", [(0, 2e-6), (10, 1e-4)])
def test_mixed_fields_and_moments_against_the_fp64_oracle(nsteps, tol):
    """The mixed form from its own initial state (exactly representable in
    FP64) against the FP64 oracle from the same state: after nsteps the six
    fields and the seven moments of each species (the output sequence
    mom_calc__accl, mom_calc__nvt, bc__mom) agree to tol of each quantity's
    maximum.  The errors are printed."""
    from wumingpic_amd import MODE_MIXED
    from wumingpic_amd.pic2d import Pic2D, weibel_charges
    nx, ny, n_ppc = 32, 32, 16
    q, r = weibel_charges(n_ppc, 0.1, 25.0)
    p = po.make_params(nx, ny, np_cap=3 * n_ppc * nx, delt=0.5, q=q, r=r)
    with Pic2D(nx, ny, q, r, delt=0.5, mode=MODE_MIXED, capacity=2 * n_ppc * nx * ny) as mix:
        mix.ic_beams(20261202, n_ppc, 0.5, 0.1, 0.01, 0.0)
        up, np2, cc = mix.download_up(p.np)
        uf = mix.download_fields()
        mix.step(nsteps)
        mix.moments()
        mom_mix = mix.download_mom()
        uf_mix = mix.download_fields()
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    if nsteps:
        o.step(up, gp, uf, cc, np2, nsteps)
    o.mom_accl(gp, up, uf, cc)
    mom = o.mom_nvt(gp, np2)
    o.bc_mom(mom)
    err = _moment_errors(mom_mix, mom)
    ferr = {}
    for c, name in enumerate(("Bx", "By", "Bz", "Ex", "Ey", "Ez")):
        scale = np.abs(uf[c]).max()
        if scale > 0:
            ferr[name] = float(np.abs(uf_mix[c] - uf[c]).max() / scale)
    print(f"mixed vs FP64 oracle after {nsteps} steps: moments {err} fields {ferr}")
    assert max(err.values()) <= tol, err
    assert not ferr or max(ferr.values()) <= max(tol, 1e-4), ferr
