"""Device adjoint of the 2D periodic loop (include/wpic_adj.h) through the
torch.autograd.Function harness: gradients of the final magnetic energy with
respect to the beam drift, the electron temperature and the amplitude of a
Bz perturbation, against central finite differences of the forward (step
1e-5, CG tolerance 1e-13 so that the forward is smooth in its inputs).
Tolerance: 1e-3 relative."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _fd(case, theta, h=1e-5):
    from wumingpic_amd.adjoint import magnetic_energy
    g = []
    for k in range(3):
        tp, tm = list(theta), list(theta)
        tp[k] += h
        tm[k] -= h
        g.append((magnetic_energy(case, *tp) - magnetic_energy(case, *tm)) / (2 * h))
    return np.array(g)


@pytest.mark.parametrize("nsteps,every", [(12, 4), (10, 20)])
def test_adjoint_gradients_match_central_differences(nsteps, every):
    from wumingpic_amd.adjoint import WeibelCase, WeibelMagneticEnergy, magnetic_energy
    case = WeibelCase(nx=32, ny=32, n_ppc=8, nsteps=nsteps, every=every)
    theta = [0.5, 0.1, 1e-3]
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    wb = WeibelMagneticEnergy.apply(t, case)
    wb.backward()
    grad = t.grad.numpy()
    wbv = float(wb.detach())
    assert abs(wbv - magnetic_energy(case, *theta)) <= 1e-12 * abs(wbv)
    fd = _fd(case, theta)
    rel = np.abs(grad - fd) / np.abs(fd)
    assert (rel <= 1e-3).all(), (grad, fd, rel)


def test_adjoint_of_the_initial_state_downloads_in_reference_layout():
    from wumingpic_amd.adjoint import WeibelCase, make_sim
    case = WeibelCase(nx=16, ny=16, n_ppc=4, nsteps=3, every=2)
    with make_sim(case, 0.5, 0.1, 1e-3) as sim:
        sim.adj_forward(case.nsteps, case.every)
        sim.adj_backward(1.0)
        lam_up, lam_uf = sim.adj_download(2 * 4 * 16)
    assert np.isfinite(lam_up).all() and np.isfinite(lam_uf).all()
    assert np.abs(lam_uf[:, 2:18, 2:18]).max() > 0
    assert np.abs(lam_uf[:, :2, :]).max() == 0 and np.abs(lam_uf[:, :, :2]).max() == 0
    assert np.abs(lam_up[2:5]).max() > 0


def test_adjoint_called_from_fortran_matches_the_harness():
    """the adjoint entry points through their ISO_C_BINDING interfaces
    (wumingpic_amd/fortran/adj_driver) give the harness's gradients"""
    import os
    import subprocess

    from wumingpic_amd.adjoint import WeibelCase, WeibelMagneticEnergy
    exe = os.path.join(os.path.dirname(__file__), "..", "wumingpic_amd", "fortran", "adj_driver")
    out = subprocess.run([exe, "32", "32", "8", "12", "4", "0.5", "0.1", "1e-3"], capture_output=True, text=True,
                         check=True, timeout=120).stdout.split()
    wb_f, g_f = float(out[0]), np.array([float(v) for v in out[1:4]])
    case = WeibelCase(nx=32, ny=32, n_ppc=8, nsteps=12, every=4)
    t = torch.tensor([0.5, 0.1, 1e-3], dtype=torch.float64, requires_grad=True)
    wb = WeibelMagneticEnergy.apply(t, case)
    wb.backward()
    assert abs(wb_f - float(wb.detach())) <= 1e-12 * abs(wb_f)
    assert np.allclose(g_f, t.grad.numpy(), rtol=1e-9, atol=0)
