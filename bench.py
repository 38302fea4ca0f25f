"""Benchmark of the MI355X PIC step loop.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--start S] [--workload weibel1024]

Workload (default): 2D Weibel beam instability, 1024 x 1024 periodic cells,
64 particles per cell per species (134,217,728 particles), mass ratio 25,
dt = 0.5, omega_pe dt = 0.05, electron beams u_x = +/-0.5 with 0.1 thermal
spread, ions 0.01; FP64 path.  One step = push + deposit + implicit field
solve + particle boundaries + row exchange + sort, fused on the device.
Inputs are resident in HBM before the timed region.

The run advances S steps (default 880) untimed, then W warmup steps, then
times exactly K steps, so with the defaults of the driver (K = 20, W = 5) the
timed window is steps 885-905 of the workload's 1000: late in the
instability, where the plasma is hotter and more particles change cell than
at the start.

Prints one JSON line: particle-pushes/s for the whole job, the roofline of the
dominant kernel (push+deposit, HIP-event timed on the library's own stream in
the timed region) and a CPU baseline: the oracle (an OpenMP C restatement of
the reference step) timed on this host's cores on a bounded sample of the same
workload.  After the timed region the same number of steps runs again without
any HIP event (ms_per_step_without_events) and once more with the events of
every phase (phase_ms_per_step).  The "mixed" object is the mixed-precision
form (int32 cell + FP32 offsets, FP32 push and deposit, FP64 field solve) run
from the same seeded state through the same steps, with its speed-up and its
field, moment and energy error against the FP64 run at the same step.  The
"traffic" of the roofline comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE
summary committed under profiles/ for this kernel, when present.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one particle through the push+deposit kernel: the
# 48-byte record (x, y, ux, uy, uz, id) read once and written once, the
# per-particle-step figure of the whole loop (the sort moves no record)
PUSH_DEPOSIT_BYTES = 96
STEP_BYTES = 96
# the kernel that implements the fused push + deposit of the FP64 fast mode
PUSH_DEPOSIT_KERNEL = "k_push_deposit_fast<16, true, 0, false>"
# mixed form: read and write the 32-byte record (cell, 2 offsets, 3 momenta, id)
PUSH_DEPOSIT_BYTES_MIXED = 64

WORKLOADS = {
    "weibel1024": dict(nx=1024, ny=1024, n_ppc=64,
                       name="2D Weibel beams 1024x1024, 64 ppc/species, FP64 path"),
    "weibel64": dict(nx=64, ny=64, n_ppc=16, name="2D Weibel beams 64x64, 16 ppc/species, FP64 path"),
    # a 3D run for reference figures (not the headline line): 3d/proj/weibel at dt = 0.5
    "weibel3d256": dict(nx=256, ny=256, nz=256, n_ppc=16,
                        name="3D Weibel beams 256x256x256, 16 ppc/species, FP64 path"),
    "weibel3d64": dict(nx=64, ny=64, nz=64, n_ppc=16, name="3D Weibel beams 64x64x64, 16 ppc/species, FP64 path"),
}
# algorithmic bytes of one 3D particle through the push or the fused fast
# push+deposit (read the 56-byte record x,y,z,u,id and write it) and through
# the separate deposit (read old and new x,y,z)
PUSH3D_BYTES = 112
DEPOSIT3D_BYTES = 48
# 3D mixed form: the 36-byte record (cell, 3 offsets, 3 momenta, id) each way
PUSH3D_BYTES_MIXED = 72
# moments (fused kick + CIC deposit): the 40-byte record without its id read
# once through the 4-byte permutation entry; nothing per particle is written
MOM_BYTES = 44
# the mixed form's: cell, two FP32 offsets and three FP32 momenta (24 bytes)
# and the permutation entry
MOM_BYTES_MIXED = 28
# 3D: the 48-byte record and the permutation entry
MOM3D_BYTES = 52


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--start", type=int, default=880,
                    help="untimed steps before the warmup (the timed window starts at step start + warmup)")
    ap.add_argument("--workload", default="weibel1024", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="fast", choices=["fast", "parity"])
    ap.add_argument("--cpu-rows", type=int, default=128, help="rows of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison run")
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


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary"""
    path = os.path.join(ROOT, "profiles", "push_deposit_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel:
        return None
    return t.get("traffic_bytes_per_launch")


def timed_run(sim, start, warmup, steps):
    """advance `start` + `warmup` steps, time exactly `steps` steps (HIP events
    of the library on: the dominant kernel's launch times), then run `steps`
    more without events and `steps` more with the events of every phase.
    Returns (timed seconds, seconds of the event-free window, per-phase ms per
    step, push_deposit (ms, launches) of the timed region)."""
    import torch
    sim.step(start + warmup)
    torch.cuda.synchronize()
    sim.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pd = sim.timing("push_deposit")
    sim.set_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    plain = time.perf_counter() - t0
    sim.set_timing(True)
    sim.step(steps)
    torch.cuda.synchronize()
    phases = {name: sim.timing(name)[0] / steps
              for name in ("push_deposit", "push", "deposit", "field", "cg", "exchange", "sort")
              if sim.timing(name)[1] > 0}
    field = sim.timing("field")
    sim.set_timing(False)
    return elapsed, plain, phases, pd, field


def moments_run(sim, npart, nbytes, reps=5):
    """the output-cadence moments (mom_calc__accl + __nvt + bc__mom fused) on
    the state the steps left, outside the timed region; returns the timing and
    the moments"""
    import torch
    sim.set_timing(True)
    for _ in range(reps):
        sim.moments()
    torch.cuda.synchronize()
    ms, n = sim.timing("moments")
    sim.set_timing(False)
    avg = ms / max(n, 1)
    return {"ms": avg, "launches": n, "bytes_per_particle": nbytes,
            "achieved_GBs": nbytes * npart / (avg / 1e3) / 1e9,
            "what": "mom_calc__accl, mom_calc__nvt and bc__mom in one pass (kick fused into the deposit)"}, \
        sim.download_mom()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


MOMENT_NAMES = ("N", "Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz")
FIELD_NAMES = ("Bx", "By", "Bz", "Ex", "Ey", "Ez")


def rel_err(a, ref):
    scale = float(abs(ref).max())
    return float(abs(a - ref).max() / scale) if scale > 0 else None


def mixed_run(w, a, npart, ref):
    """the mixed-precision form through the same steps from the same seed;
    its errors against the FP64 run at the same step"""
    import numpy as np

    from wumingpic_amd import MODE_MIXED

    if "nz" in w:
        from wumingpic_amd.pic3d import weibel_beams_3d
        sim = weibel_beams_3d(w["nx"], w["ny"], w["nz"], w["n_ppc"], mode=MODE_MIXED,
                              capacity=int(w["n_ppc"] * w["nx"] * w["ny"] * w["nz"] * 1.05))
        nbytes = PUSH3D_BYTES_MIXED
    else:
        from wumingpic_amd.pic2d import weibel_beams_2d
        sim = weibel_beams_2d(w["nx"], w["ny"], w["n_ppc"], mode=MODE_MIXED)
        nbytes = PUSH_DEPOSIT_BYTES_MIXED
    elapsed, plain, phases, pd, _ = timed_run(sim, a.start, a.warmup, a.steps)
    log(f"mixed form: {a.start} + {a.warmup} untimed + {a.steps} timed steps: {elapsed:.3f} s")
    energy = sim.energy()
    uf = sim.download_fields()
    mom = mom_t = None
    if "nz" not in w:
        from wumingpic_amd import WpicError
        try:
            mom_t, mom = moments_run(sim, npart, MOM_BYTES_MIXED)
        except WpicError as e:  # a library without the mixed-form moments
            log(f"mixed-form moments unavailable: {e}")
    sim.close()
    pd_avg_s = pd[0] / max(pd[1], 1) / 1e3
    out = {
        "value": npart * a.steps / elapsed,
        "unit": "particle-pushes/s",
        "ms_per_step": elapsed / a.steps * 1e3,
        "ms_per_step_without_events": plain / a.steps * 1e3,
        "speedup_vs_f64": ref["elapsed"] / elapsed,
        "speedup_vs_f64_without_events": ref["plain"] / plain,
        "dtype": "int32 cell + f32 offsets/momenta; f32 push/deposit; f64 field solve and current sum",
        "push_deposit": {"avg_launch_ms": pd_avg_s * 1e3, "bytes_per_particle": nbytes,
                         "achieved_GBs": nbytes * npart / pd_avg_s / 1e9},
        "phase_ms_per_step": phases,
        "compared_at_step": ref["step"],
        "field_rel_err_vs_f64": {nm: rel_err(uf[c], ref["uf"][c]) for c, nm in enumerate(FIELD_NAMES)},
        "energy_rel_err_vs_f64": {k: float(abs(energy[i] - ref["energy"][i]) / abs(ref["energy"][i]))
                                  for i, k in enumerate(("kinetic", "E", "B"))},
    }
    if mom_t is not None:
        out["moments"] = mom_t
    if mom is not None and ref.get("mom") is not None:
        out["moment_rel_err_vs_f64"] = {f"{nm}[{sp}]": rel_err(mom[k, :, :, sp], ref["mom"][k, :, :, sp])
                                        for sp in range(mom.shape[3]) for k, nm in enumerate(MOMENT_NAMES)}
    return out


def main3d(a, w):
    """3D figures: pushes/s and the per-phase times of the fused 3D step"""
    from wumingpic_amd import MODE_FAST, MODE_PARITY
    from wumingpic_amd.pic3d import weibel_beams_3d

    mode = MODE_PARITY if a.mode == "parity" else MODE_FAST
    sim = weibel_beams_3d(w["nx"], w["ny"], w["nz"], w["n_ppc"], mode=mode,
                          capacity=int(w["n_ppc"] * w["nx"] * w["ny"] * w["nz"] * 1.05))
    npart = int(sim.counts().sum())
    log(f"initial state ready: {w['name']}, {npart} particles")
    elapsed, plain, phases, pd, _ = timed_run(sim, a.start, a.warmup, a.steps)
    log(f"{a.start} + {a.warmup} untimed + {a.steps} timed steps: {elapsed:.3f} s")
    cg = sim.cg_iters()
    moments, _ = moments_run(sim, npart, MOM3D_BYTES)
    energy64 = sim.energy()
    uf64 = sim.download_fields() if not a.no_mixed else None
    sim.close()
    rates = {}
    for name, nbytes in (("push_deposit", PUSH3D_BYTES), ("push", PUSH3D_BYTES), ("deposit", DEPOSIT3D_BYTES)):
        if name in phases:
            rates[name + "_GBs"] = nbytes * npart / (phases[name] / 1e3) / 1e9
    out = {
        "metric": "particle-pushes/s (push+deposit+sort) per GPU, 3D",
        "value": npart * a.steps / elapsed, "unit": "particle-pushes/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": elapsed / a.steps * 1e3,
        "ms_per_step_without_events": plain / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded counter-based initial state, generated on the device)",
        "config": {"workload": w["name"], "particles": npart, "mode": a.mode,
                   "timed_steps": [a.start + a.warmup, a.start + a.warmup + a.steps]},
        "phase_ms_per_step": phases,
        **rates,
        "cg_iters_last_step": list(cg),
        "moments": moments,
    }
    if not a.no_mixed:
        ref = {"elapsed": elapsed, "plain": plain, "uf": uf64, "energy": energy64,
               "step": a.start + a.warmup + 3 * a.steps}
        out["mixed"] = mixed_run(w, a, npart, ref)
    print(json.dumps(out))


def main():
    a = parse()
    if a.probe == "copy":
        copy_probe(a.probe_mib)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for {a.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401

    from wumingpic_amd import MODE_FAST, MODE_PARITY
    from wumingpic_amd.pic2d import weibel_beams_2d

    w = WORKLOADS[a.workload]
    if "nz" in w:
        main3d(a, w)
        return
    nx, ny, n_ppc = w["nx"], w["ny"], w["n_ppc"]
    mode = MODE_PARITY if a.mode == "parity" else MODE_FAST
    sim = weibel_beams_2d(nx, ny, n_ppc, mode=mode)
    npart = int(sim.counts().sum())
    ncell = nx * ny
    log(f"initial state ready: {w['name']}, {npart} particles")
    elapsed, plain, phases, (pd_ms, pd_n), (fld_ms, fld_n) = timed_run(sim, a.start, a.warmup, a.steps)
    log(f"{a.start} + {a.warmup} untimed + {a.steps} timed steps: {elapsed:.3f} s")
    cg_iters = sim.cg_iters()
    moments, mom64 = moments_run(sim, npart, MOM_BYTES)
    energy = sim.energy()
    uf64 = sim.download_fields()
    sim.close()
    mixed = None
    if not a.no_mixed:
        ref = {"elapsed": elapsed, "plain": plain, "uf": uf64, "energy": energy, "mom": mom64,
               "step": a.start + a.warmup + 3 * a.steps}
        mixed = mixed_run(w, a, npart, ref)

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
        "ms_per_step_without_events": plain / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based initial state, generated on the device)",
        "config": {"workload": w["name"], "nx": nx, "ny": ny, "particles_per_cell_per_species": n_ppc,
                   "particles": npart, "mode": a.mode, "parallelism": "single GPU",
                   "timed_steps": [a.start + a.warmup, a.start + a.warmup + a.steps]},
        "roofline": {"kernel": "push_deposit", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(PUSH_DEPOSIT_KERNEL),
                     "bytes_per_particle": PUSH_DEPOSIT_BYTES, "avg_launch_ms": pd_avg_s * 1e3,
                     "launches": pd_n},
        "step_bandwidth_GBs": STEP_BYTES * npart * a.steps / elapsed / 1e9,
        "cell_updates_per_s": ncell * fld_n / (fld_ms / 1e3) if fld_ms > 0 else None,
        "field_ms_per_step": fld_ms / max(fld_n, 1),
        "phase_ms_per_step": phases,
        "cg_iters_last_step": list(cg_iters),
        "moments": moments,
        "energy_final": {"kinetic": energy[0], "E": energy[1], "B": energy[2]},
        "cpu_baseline": None,
        "mixed": mixed,
    }
    if not a.no_cpu_baseline:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(w, min(a.cpu_rows, ny), a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
