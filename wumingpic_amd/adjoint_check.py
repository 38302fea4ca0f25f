"""Gradient check of the device adjoint on the 2D Weibel adjoint case:
256 x 256 periodic cells, 16 particles per cell per species, 200 steps in
FP64, checkpoints every 20 steps; gradients of the final magnetic energy with
respect to the beam drift (0.5), the electron thermal spread (0.1) and the
amplitude (1e-3) of a sinusoidal Bz perturbation, against central finite
differences with step 1e-5.

    python -m wumingpic_amd.adjoint_check [--nx 256 --ny 256 --n-ppc 16 --steps 200 --every 20]

Prints one JSON line.
"""
import argparse
import json
import time

import numpy as np
import torch

from .adjoint import WeibelCase, WeibelMagneticEnergy, magnetic_energy


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--n-ppc", type=int, default=16)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--h", type=float, default=1e-5)
    a = ap.parse_args()
    case = WeibelCase(nx=a.nx, ny=a.ny, n_ppc=a.n_ppc, nsteps=a.steps, every=a.every)
    theta = [0.5, 0.1, 1e-3]
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    wb = WeibelMagneticEnergy.apply(t, case)
    t1 = time.perf_counter()
    wb.backward()
    t2 = time.perf_counter()
    grad = t.grad.numpy()
    fd = []
    for k in range(3):
        tp, tm = list(theta), list(theta)
        tp[k] += a.h
        tm[k] -= a.h
        fd.append((magnetic_energy(case, *tp) - magnetic_energy(case, *tm)) / (2 * a.h))
    t3 = time.perf_counter()
    fd = np.array(fd)
    names = ("udrift", "uth_e", "bz_amplitude")
    print(json.dumps({
        "case": f"2D Weibel {a.nx}x{a.ny}, {a.n_ppc} ppc/species, {a.steps} steps FP64, checkpoint every {a.every}",
        "W_B": float(wb.detach()),
        "adjoint": dict(zip(names, map(float, grad))),
        "central_fd": dict(zip(names, map(float, fd))),
        "rel_err": dict(zip(names, map(float, np.abs(grad - fd) / np.abs(fd)))),
        "fd_step": a.h,
        "seconds": {"forward_recorded": t1 - t0, "backward": t2 - t1, "six_fd_forwards": t3 - t2},
    }))


if __name__ == "__main__":
    main()
