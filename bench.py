"""Benchmark of the MI355X PIC step loop.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload weibel1024]

Workload (default): 2D Weibel beam instability, 1024 x 1024 periodic cells,
64 particles per cell per species (134,217,728 particles), mass ratio 25,
dt = 0.5, omega_pe dt = 0.05, electron beams u_x = +/-0.5 with 0.1 thermal
spread, ions 0.01; FP64 path.  One step = push + deposit + implicit field
solve + particle boundaries + row exchange + sort, fused on the device.
Inputs are resident in HBM before the timed region.

Prints one JSON line: particle-pushes/s for the whole job, the roofline of the
dominant kernel (push+deposit, HIP-event timed on the library's own stream)
and a CPU baseline: the oracle (an OpenMP C restatement of the reference step)
timed on this host's cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one particle through the push+deposit kernel: read
# x, y, ux, uy, uz and write them back (5 doubles each way)
PUSH_DEPOSIT_BYTES = 80
# one particle-step of the whole loop: the 48-byte record read once and
# written once to its sorted slot
STEP_BYTES = 96

WORKLOADS = {
    "weibel1024": dict(nx=1024, ny=1024, n_ppc=64,
                       name="2D Weibel beams 1024x1024, 64 ppc/species, FP64 path"),
    "weibel64": dict(nx=64, ny=64, n_ppc=16, name="2D Weibel beams 64x64, 16 ppc/species, FP64 path"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="weibel1024", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="fast", choices=["fast", "parity"])
    ap.add_argument("--cpu-rows", type=int, default=128, help="rows of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe", choices=["copy"], default=None,
                    help="run a plain device copy of --probe-mib MiB (read) instead, to calibrate "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--probe-mib", type=int, default=1024)
    return ap.parse_args()


def copy_probe(mib):
    """one elementwise kernel that reads mib MiB and writes mib MiB"""
    import torch
    n = mib * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "copy", "read_mib": mib, "write_mib": mib, "launches": 3}))


def cpu_baseline(w, rows, steps):
    """The oracle (C, OpenMP) on a sample of the same workload: same row width,
    particles per cell and physics, `rows` periodic rows instead of ny."""
    import numpy as np

    from oracle import pyoracle as po
    from wumingpic_amd.pic2d import weibel_beams_2d, weibel_charges

    nx, n_ppc = w["nx"], w["n_ppc"]
    q, r = weibel_charges(n_ppc, 0.1, 25.0)
    p = po.make_params(nx, rows, np_cap=2 * n_ppc * nx, delt=0.5, q=q, r=r)
    with weibel_beams_2d(nx, rows, n_ppc, capacity=n_ppc * nx * rows) as sim:
        up, np2, cc = sim.download_up(p.np)
        uf = sim.download_fields()
    cores = len(os.sched_getaffinity(0))
    o = po.Oracle2D(p, threads=cores)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 1)  # first touch of the work arrays
    t0 = time.perf_counter()
    o.step(up, gp, uf, cc, np2, steps)
    dt = time.perf_counter() - t0
    n = int(np2.sum())
    return {"value": n * steps / dt, "unit": "particle-pushes/s", "cores": cores, "kind": "port",
            "sample": f"{nx}x{rows} periodic cells (same row width, {n_ppc} ppc/species, physics and dt "
                      f"as the GPU workload), {n} particles, {steps} steps after 1 untimed, "
                      f"{dt:.2f} s"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.probe == "copy":
        copy_probe(a.probe_mib)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for {a.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)
    import torch

    from wumingpic_amd import MODE_FAST, MODE_PARITY
    from wumingpic_amd.pic2d import weibel_beams_2d

    w = WORKLOADS[a.workload]
    nx, ny, n_ppc = w["nx"], w["ny"], w["n_ppc"]
    mode = MODE_PARITY if a.mode == "parity" else MODE_FAST
    sim = weibel_beams_2d(nx, ny, n_ppc, mode=mode)
    npart = int(sim.counts().sum())
    ncell = nx * ny
    log(f"initial state ready: {w['name']}, {npart} particles")
    sim.step(a.warmup)
    torch.cuda.synchronize()
    log(f"{a.warmup} warmup steps done")
    sim.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(a.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    log(f"{a.steps} timed steps in {elapsed:.3f} s")
    pd_ms, pd_n = sim.timing("push_deposit")
    fld_ms, fld_n = sim.timing("field")
    phases = {name: sim.timing(name)[0] / a.steps for name in ("push_deposit", "field", "cg", "exchange", "sort")}
    cg_iters = sim.cg_iters()
    sim.set_timing(False)
    energy = sim.energy()
    sim.close()

    pd_avg_s = pd_ms / max(pd_n, 1) / 1e3
    achieved = PUSH_DEPOSIT_BYTES * npart / pd_avg_s / 1e9
    out = {
        "metric": "particle-pushes/s (push+deposit+sort) per GPU",
        "value": npart * a.steps / elapsed,
        "unit": "particle-pushes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based initial state, generated on the device)",
        "config": {"workload": w["name"], "nx": nx, "ny": ny, "particles_per_cell_per_species": n_ppc,
                   "particles": npart, "mode": a.mode, "parallelism": "single GPU"},
        "roofline": {"kernel": "push_deposit", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_particle": PUSH_DEPOSIT_BYTES, "avg_launch_ms": pd_avg_s * 1e3,
                     "launches": pd_n},
        "step_bandwidth_GBs": STEP_BYTES * npart * a.steps / elapsed / 1e9,
        "cell_updates_per_s": ncell * fld_n / (fld_ms / 1e3) if fld_ms > 0 else None,
        "field_ms_per_step": fld_ms / max(fld_n, 1),
        "phase_ms_per_step": phases,
        "cg_iters_last_step": list(cg_iters),
        "energy_final": {"kinetic": energy[0], "E": energy[1], "B": energy[2]},
        "cpu_baseline": None,
    }
    if not a.no_cpu_baseline:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(w, min(a.cpu_rows, ny), a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
