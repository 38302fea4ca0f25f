"""One Fourier mode of the staggered fields through the time-centred Maxwell
system: the closed form the field-update tests (oracle and device) check the
step against.

The update of field__fdtd_i (2d/common/field.f90:124-184,
3d/common/field.f90:127-206) is Maxwell's equations on the staggered grid,
time-centred with weight g (gfac):
    B' = B - c dt curl(g E' + (1 - g) E)
    E' = E + c dt curl(g B' + (1 - g) B) - 4 pi dt J
Each difference is centred between two staggered nodes, so on exp(i k.x)
each d/dx acts as i K_x, K_x = 2 sin(k_x dx / 2) / dx.  Offsets of the
components (Bx, By, Bz, Ex, Ey, Ez) within a cell, in units of dx; the
currents sit on their E components.
"""
import numpy as np

OFF2 = np.array([[0.5, 0.0], [0.0, 0.5], [0.0, 0.0], [0.0, 0.5], [0.5, 0.0], [0.5, 0.5]])
OFF3 = np.array([[0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5],
                 [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])


def cross_matrix(v):
    """M with M @ w = v x w"""
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], dtype=complex)


def mode_step(K, B, E, J, c, dt, g):
    """one step of the time-centred Maxwell system for amplitudes of
    exp(i k.x); K the difference symbols (real), J the current amplitude"""
    C = cross_matrix(1j * np.asarray(K, dtype=complex))
    I3 = np.eye(3)
    A = np.block([[I3, c * dt * g * C], [-c * dt * g * C, I3]])
    rhs = np.concatenate([B - c * dt * (1 - g) * (C @ E), E + c * dt * (1 - g) * (C @ B) - 4 * np.pi * dt * J])
    x = np.linalg.solve(A, rhs)
    return x[:3], x[3:]


def plane(amp, k, off, shape, lo):
    """Re(amp exp(i k.(n + off))) on an index box starting at lo (per axis)"""
    grids = np.meshgrid(*[np.arange(n) + l for n, l in zip(shape, lo)], indexing="ij")
    ph = sum(kk * (gg + o) for kk, gg, o in zip(k, grids, off))
    return np.real(amp * np.exp(1j * ph))


def fields_of(B, E, k, offs, shape, lo):
    F = np.concatenate([B, E])
    return np.asfortranarray(np.stack([plane(F[c], k, offs[c], shape, lo) for c in range(6)]))


def current_of(J, k, offs, shape, lo):
    return np.asfortranarray(np.stack([plane(J[c], k, offs[3 + c], shape, lo) for c in range(3)]))


def mode_amplitudes(K, rng):
    """random complex B (transverse to K: div B = 0), E and J amplitudes"""
    B = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    B -= K * (K @ B) / (K @ K)
    E = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    J = 0.05 * (rng.standard_normal(3) + 1j * rng.standard_normal(3))
    return B, E, J
