"""The reference apps' initial states and the problem driver (wumingpic_amd/app/wpic_app2d over
include/wpic_app.h) on the GPU: the Weibel app's lattice state and the reconnection app's
Harris sheet against their formulas (2d/proj/weibel/app.f90:366-438,
2d/proj/reconnection/app.f90:260-452), and whole driver runs in parity mode against the oracle
run from the same initial state: the moment file, the final restart file and energy.dat; a run
stopped at the elapsed-time limit and restarted from its snapshot ends bit-identical to an
uninterrupted run."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "wumingpic_amd", "app", "wpic_app2d")
PI = 3.141592653589793


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def weibel_params(n_ppc, nx, ny, mass_ratio, sigma_e, omega_pe):
    """2d/proj/weibel/app.f90:300-318"""
    wpe = omega_pe
    wge = omega_pe * np.sqrt(sigma_e)
    wpi = wpe / np.sqrt(mass_ratio)
    wgi = wge / mass_ratio
    r = (mass_ratio, 1.0)
    q = (+np.sqrt(r[0] / (4 * PI * n_ppc / 1.0)) * wpi, -np.sqrt(r[1] / (4 * PI * n_ppc / 1.0)) * wpe)
    b0 = r[0] * 1.0 / q[0] * wgi
    return q, r, b0


def harris_params(nx, ny, mass_ratio, alpha, rtemp, lcs, nbg, ncs):
    """2d/proj/reconnection/app.f90:281-305"""
    r = (mass_ratio, 1.0)
    c = 1.0
    vte = np.sqrt(rtemp) * c / (np.sqrt(1 + rtemp) * alpha)
    vti = vte * np.sqrt(r[1] / r[0]) / np.sqrt(rtemp)
    wpe = vte / 1.0 / np.sqrt(2.0)
    wpi = wpe * np.sqrt(r[1] / r[0])
    wge = wpe / alpha
    wgi = wge / mass_ratio
    n0 = nbg + ncs
    q = (+np.sqrt(r[0] / (4 * PI * n0)) * wpi, -np.sqrt(r[1] / (4 * PI * n0)) * wpe)
    b0 = r[0] * c / q[0] * wgi
    nxgs, nxge, nygs, nyge = 2, 1 + nx, 2, 1 + ny
    return dict(q=q, r=r, b0=b0, vte=vte, vti=vti, x0=0.5 * (nxge + nxgs), y0=0.5 * (nyge - nygs),
                lcs=lcs * c / wpi, np2=int(nbg * (nxge - nxgs) + ncs * 2 * (lcs * c / wpi)),
                ibg=nbg * (nxge - nxgs - 2))


def _pic(nx, ny, q, r, bc, mode, cap, delt):
    from wumingpic_amd.pic2d import Pic2D
    return Pic2D(nx, ny, list(q), list(r), delt=delt, bc=bc, mode=mode, capacity=cap)


def test_weibel_initial_state_matches_its_formulas():
    from wumingpic_amd import MODE_PARITY
    nx, ny, n0 = 24, 10, 6
    q, r, b0 = weibel_params(n0, nx, ny, 25.0, 0.0, 0.1)
    print("weibel initial state", flush=True)
    with _pic(nx, ny, q, r, po.BC_PERIODIC, MODE_PARITY, n0 * nx * ny, 1.0) as sim:
        sim.ic_weibel(7, n0, 0.02, 0.1, 5.0, b0)
        up, np2, cc = sim.download_up(n0 * nx)
        uf = sim.download_fields()
    assert np.all(np2 == n0 * nx)
    assert np.all(cc == (np.arange(nx + 1) * n0)[:, None, None])
    ii = np.arange(1, n0 * nx + 1)
    x = (2 + nx * (ii - 0.5) / (n0 * nx)) * 1.0  # app.f90:397
    for s in range(2):
        for j in range(ny):
            assert np.array_equal(up[0, :, j, s], x)
            y = up[1, :, j, s]
            assert np.all((y >= 2 + j) & (y < 3 + j))
            ids = up[5, :, j, s].view(np.int64)
            assert np.array_equal(ids, -(j * n0 * nx + ii))
    assert np.array_equal(up[1, :, :, 0], up[1, :, :, 1])  # co-located species
    for s, sd in ((0, 0.02), (1, 0.1)):
        u = up[2:5, :, :, s].reshape(3, -1)
        assert np.allclose(u.std(axis=1), [sd, sd, 5.0 * sd], rtol=0.06)
        assert np.all(np.abs(u.mean(axis=1)) < 4 * np.array([sd, sd, 5 * sd]) / np.sqrt(u.shape[1]))
    assert np.all(uf[2] == b0) and np.all(uf[[0, 1, 3, 4, 5]] == 0)


def test_harris_initial_state_matches_its_formulas():
    from wumingpic_amd import MODE_PARITY
    nx, ny = 64, 24
    hp = harris_params(nx, ny, 25.0, 2.0, 0.2, 0.5, 8, 40)
    e1 = 0.12
    print("harris initial state", flush=True)
    with _pic(nx, ny, hp["q"], hp["r"], po.BC_WALLX, MODE_PARITY, hp["np2"] * ny, 0.5) as sim:
        sim.ic_harris(11, 8, 40, hp["lcs"], hp["x0"], hp["y0"], hp["b0"], e1, 0.2, hp["vti"], hp["vte"])
        up, np2, cc = sim.download_up(2 * hp["np2"])
        uf = sim.download_fields()
    assert np.all(np2 == hp["np2"])
    # the field at the nodes (app.f90:397-412)
    i = np.arange(0, nx + 4)[:, None] * 1.0
    j = np.arange(0, ny + 4)[None, :] * 1.0
    x0, y0, lcs, b0 = hp["x0"], hp["y0"], hp["lcs"], hp["b0"]
    g = np.exp(-((i - x0) ** 2 + (j - y0) ** 2) / (2 * lcs) ** 2)
    bx = e1 * b0 * ((j - y0) / lcs) * g
    by = b0 * np.tanh((i - x0) / lcs) - e1 * b0 * ((i - x0) / lcs) * g
    assert np.allclose(uf[0], bx, rtol=1e-13, atol=1e-15 * abs(b0))
    assert np.allclose(uf[1], by, rtol=1e-13, atol=1e-15 * abs(b0))
    assert np.all(uf[2:] == 0)
    # sorted rows, positions inside the walls and the row
    for s in range(2):
        for jj in range(ny):
            n = np2[jj, s]
            x = up[0, :n, jj, s]
            assert np.all(np.diff(np.floor(x)) >= 0)
            assert np.all((x >= 3) & (x < nx + 1)) and np.all(np.floor(up[1, :n, jj, s]) == 2 + jj)
    # the density profile: nbg per unit length uniform on [3, nx) plus the sheet's np2 - ibg
    # particles with the truncated cosh^-2 profile (app.f90:383-403), per bin of 4 cells over rows
    xs = up[0, :hp["np2"], :, 0].ravel()
    edges = np.arange(4.0, nx - 3, 4.0)
    hist, _ = np.histogram(xs, bins=edges)
    L = nx - 3
    ncs_row = hp["np2"] - hp["ibg"]
    sheet = ncs_row * (np.tanh((edges[1:] - x0) / lcs) - np.tanh((edges[:-1] - x0) / lcs)) / (
        2 * np.tanh(0.5 * L / lcs))
    expect = ny * (8 * 4.0 + sheet)
    assert np.all(np.abs(hist - expect) < 5 * np.sqrt(expect)), (hist, expect)
    # the drift that carries jz at the sheet centre: mean uz of ions > 0, electrons < 0
    centre = np.abs(up[0, :hp["np2"], :, :] - x0) < 0.5 * lcs
    assert up[4, :hp["np2"], :, 0][centre[:, :, 0]].mean() > 0
    assert up[4, :hp["np2"], :, 1][centre[:, :, 1]].mean() < 0


def _run_app(cfg, d, mode="parity"):
    path = os.path.join(d, "config.json")
    with open(path, "w") as f:
        json.dump(cfg, f)
    r = subprocess.run([APP, path, mode], capture_output=True, text=True, timeout=120, cwd=d)
    assert r.returncode == 0, r.stdout + r.stderr
    return path


def _read_restart(d, name, np_cap, ny, nx):
    meta = json.load(open(os.path.join(d, name + ".json")))
    raw = open(os.path.join(d, name + ".raw"), "rb").read()
    ds = meta["dataset"]
    up = np.frombuffer(raw, dtype="<f8", count=6 * np_cap * ny * 2, offset=ds["up"]["offset"]).reshape(
        (6, np_cap, ny, 2), order="F").copy(order="F")
    np2 = np.frombuffer(raw, dtype="<i4", count=ny * 2, offset=ds["np2"]["offset"]).reshape((ny, 2), order="F")
    uf = np.frombuffer(raw, dtype="<f8", count=6 * (nx + 4) * (ny + 4), offset=ds["uf"]["offset"]).reshape(
        (6, nx + 4, ny + 4), order="F").copy(order="F")
    return meta, up, np2.copy(order="F"), uf


def _same_particles(ua, na, ub, nb):
    assert np.array_equal(na, nb)
    for s in range(na.shape[1]):
        for j in range(na.shape[0]):
            n = na[j, s]
            assert np.array_equal(_bits(ua[:, :n, j, s]), _bits(ub[:, :n, j, s])), (s, j)


WEIBEL = {"num_process": 1, "n_ppc": 6, "n_x": 24, "n_y": 16, "mass_ratio": 25.0, "sigma_e": 0.0,
          "omega_pe": 0.1, "v_the": 0.1, "v_thi": 0.02, "t_ani": 5.0}
RECON = {"num_process": 1, "n_x": 64, "n_y": 16, "mass_ratio": 4.0, "alpha": 1.0, "rtemp": 0.2,
         "lcs": 0.5, "nbg": 3, "ncs": 9}


def _setup(problem, seed):
    """the app's derived parameters and the same initial state through the library"""
    from wumingpic_amd import MODE_PARITY
    par = dict(WEIBEL if problem == "weibel" else RECON, seed=seed)
    nx, ny = par["n_x"], par["n_y"]
    if problem == "weibel":
        q, r, b0 = weibel_params(6, nx, ny, 25.0, 0.0, 0.1)
        np_cap, delt, bc, cap = 6 * nx * 5, 1.0, po.BC_PERIODIC, 6 * nx * ny
    else:
        hp = harris_params(nx, ny, 4.0, 1.0, 0.2, 0.5, 3, 9)
        q, r = hp["q"], hp["r"]
        np_cap, delt, bc, cap = 12 * nx, 0.5, po.BC_WALLX, hp["np2"] * ny
    with _pic(nx, ny, q, r, bc, MODE_PARITY, cap, delt) as sim:
        if problem == "weibel":
            sim.ic_weibel(seed, 6, 0.02, 0.1, 5.0, b0)
        else:
            sim.ic_harris(seed, 3, 9, hp["lcs"], hp["x0"], hp["y0"], hp["b0"], 0.12, 0.2, hp["vti"], hp["vte"])
        up, np2, cc = sim.download_up(np_cap)
        uf = sim.download_fields()
    p = po.make_params(nx, ny, np_cap=np_cap, delt=delt, q=q, r=r, bc=bc)
    return par, p, (q, r, delt), (up, np2, cc, uf)


@pytest.mark.parametrize("problem", ["weibel", "reconnection"])
def test_driver_run_matches_the_oracle(problem):
    """the app loop in parity mode (weibel/app.f90:99-157, reconnection/app.f90:99-148): the
    particle and moment files of the last step equal the ones the oracle's run from the same initial
    state gives, byte for byte; energy.dat has the reference's lines at the moment cadence; the
    Weibel app leaves the final state as restart max_it + 1 and the reconnection app does not"""
    from wumingpic_amd.paraio import ParaIO
    steps = 6
    par, p, (q, r, delt), (up, np2, cc, uf) = _setup(problem, 5)
    nx, ny, np_cap = par["n_x"], par["n_y"], p.np
    cfg = {"config": {"verbose": 0, "datadir": ".", "max_elapsed": 1.0e6, "max_it": steps, "intvl_ptcl": steps,
                      "intvl_mom": 3, "intvl_orb": 100}, "parameter": par}
    print(f"driver run {problem}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        _run_app(cfg, d)
        files = {n: open(os.path.join(d, n), "rb").read() for n in os.listdir(d) if n.endswith((".json", ".raw"))}
        edat = open(os.path.join(d, "energy.dat")).read().splitlines()
        final = "%07d_restart" % (steps + 1)
        restart = _read_restart(d, final, np_cap, ny, nx) if problem == "weibel" else None
    assert ("%07d_restart.json" % (steps + 1) in files) == (problem == "weibel")
    assert "init_param.json" in files and "init_param.raw" in files
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, steps)
    if restart is not None:
        meta, upf, np2f, uff = restart
        _same_particles(upf, np2f, up, np2)
        assert np.array_equal(_bits(uff), _bits(uf))
        assert meta["attribute"]["it"]["data"] == steps + 1
    with tempfile.TemporaryDirectory() as t:
        io = ParaIO(ndim=6, np_cap=np_cap, nsp=2, nxgs=2, nxge=nx + 1, nygs=2, nyge=ny + 1, delx=1.0, delt=delt,
                    c=1.0, q=list(q), r=list(r), dir=t + "/")
        io.ptcl(up, uf, np2, steps)
        io.param(*_param_args(problem, par), "init_param")
        # the moments of the last step (app.f90:121-124)
        o.mom_accl(gp, up, uf, cc)
        mom = o.mom_nvt(gp, np2)
        o.bc_mom(mom)
        io.mom(mom, uf, steps)
        io.close()
        for name in ("%07d_ptcl" % steps, "%07d_mom" % steps, "init_param"):
            for ext in (".json", ".raw"):
                mine = open(os.path.join(t, name + ext), "rb").read()
                assert files[name + ext] == mine, name + ext
    # energy.dat: steps 0, 3 and 6 as '(f8.2, 4(1x, e12.5))'
    assert len(edat) == 3
    line = edat[-1]
    assert len(line) == 8 + 4 * 13
    t_, kin, ee, eb, tot = (float(v) for v in line.split())
    e = o.energy(up, uf, np2)
    assert t_ == steps * delt
    assert np.allclose([kin, ee, eb], e, rtol=1e-4) and np.isclose(tot, e.sum(), rtol=1e-4)


def _param_args(problem, par):
    """io__param's arguments as the apps derive them (weibel/app.f90:300-318,
    reconnection/app.f90:281-305)"""
    if problem == "weibel":
        wpe = par["omega_pe"]
        wge = wpe * np.sqrt(par["sigma_e"])
        return (par["n_ppc"], wpe, wpe / np.sqrt(par["mass_ratio"]), wge, wge / par["mass_ratio"],
                par["v_thi"], par["v_the"])
    hp = harris_params(par["n_x"], par["n_y"], par["mass_ratio"], par["alpha"], par["rtemp"], par["lcs"],
                       par["nbg"], par["ncs"])
    wpe = hp["vte"] / np.sqrt(2.0)
    wpi = wpe * np.sqrt(1.0 / par["mass_ratio"])
    wge = wpe / par["alpha"]
    return (par["nbg"] + par["ncs"], wpe, wpi, wge, wge / par["mass_ratio"], hp["vti"], hp["vte"])


def test_driver_restart_at_the_elapsed_limit():
    """max_elapsed = 0 stops after the first step with a snapshot and the config pointing at it
    (app.f90:128-143); running the config again reads the snapshot, sorts it and finishes the run
    with the field solver's state fresh, as the reference's restart does (its CG starts from df = 0
    then, 2d/common/field.f90:98-110): both snapshots equal the oracle's"""
    steps = 5
    par, p, _, (up, np2, cc, uf) = _setup("weibel", 9)
    nx, ny, np_cap = par["n_x"], par["n_y"], p.np
    base = {"verbose": 0, "datadir": ".", "max_it": steps, "intvl_ptcl": 100, "intvl_mom": 100, "intvl_orb": 100}
    print("driver restart", flush=True)
    with tempfile.TemporaryDirectory() as b:
        path = _run_app({"config": dict(base, max_elapsed=0.0), "parameter": par}, b)
        cfg = json.load(open(path))
        assert cfg["config"]["restart_file"] == "0000001_restart"
        assert cfg["parameter"]["n_x"] == nx
        snap1 = _read_restart(b, "0000001_restart", np_cap, ny, nx)
        cfg["config"]["max_elapsed"] = 1.0e6
        with open(path, "w") as f:
            json.dump(cfg, f)
        r = subprocess.run([APP, path, "parity"], capture_output=True, text=True, timeout=120, cwd=b)
        assert r.returncode == 0, r.stdout + r.stderr
        snapn = _read_restart(b, "%07d_restart" % (steps + 1), np_cap, ny, nx)
    # the first step from the initial state
    o = po.Oracle2D(p)
    gp = np.zeros_like(up, order="F")
    o.step(up, gp, uf, cc, np2, 1)
    _same_particles(snap1[1], snap1[2], up, np2)
    assert np.array_equal(_bits(snap1[3]), _bits(uf))
    # the restart: io__input into gp, sort__bucket(up, gp, ...), steps 2..max_it with a fresh field state
    o2 = po.Oracle2D(p)
    gp2 = snap1[1].copy(order="F")
    np2r = snap1[2].copy(order="F")
    up2 = np.zeros_like(gp2, order="F")
    cc2 = np.zeros((nx + 1, ny, 2), dtype=np.int32, order="F")
    o2.sort_bucket(up2, gp2, cc2, np2r)
    ufr = snap1[3].copy(order="F")
    o2.step(up2, gp2, ufr, cc2, np2r, steps - 1)
    _same_particles(snapn[1], snapn[2], up2, np2r)
    assert np.array_equal(_bits(snapn[3]), _bits(ufr))


def test_shock_initial_state_matches_its_formulas():
    """2d/proj/shock/app.f90:405-470 with the box fully expanded (n_x_ini = n_x - 1)"""
    from wumingpic_amd import MODE_PARITY
    nx, ny, n0 = 32, 6, 4
    u0 = -0.3
    v0 = u0 / np.sqrt(1 + u0 * u0)
    b0, th, ph, ld = 2.0, np.pi / 3, np.pi / 2, 6.0
    np2_row = n0 * (nx - 2)
    print("shock initial state", flush=True)
    with _pic(nx, ny, (0.01, -0.01), (25.0, 1.0), po.BC_SHOCK, MODE_PARITY, np2_row * ny, 1.0) as sim:
        sim.ic_shock(4, n0, v0, 0.01, 0.05, ld, b0, th, ph)
        up, np2, cc = sim.download_up(5 * n0 * nx)
        uf = sim.download_fields()
    assert np.all(np2 == np2_row)
    ii = np.arange(1, np2_row + 1)
    x = (2 + (nx - 1) * (ii - 0.5) / np2_row) * 1.0
    for s in range(2):
        for j in range(ny):
            assert np.array_equal(up[0, :np2_row, j, s], x)  # the lattice is already in cell order
            assert np.all(np.floor(up[1, :np2_row, j, s]) == 2 + j)
            ids = -up[5, :np2_row, j, s].view(np.int64)
            assert np.array_equal(np.sort(ids), j * np2_row + ii)
    vprof = lambda xx: 0.5 * v0 * (1 + np.tanh((xx - (ld + 2)) / (0.1 * ld)))  # noqa: E731
    i = np.arange(0, nx + 4)[:, None] * 1.0
    assert np.all(uf[0] == b0 * np.cos(th)) and np.all(uf[3] == 0)
    assert np.allclose(uf[4], vprof(i) * uf[2], rtol=1e-14, atol=1e-16)
    assert np.allclose(uf[5], -vprof(i) * uf[1], rtol=1e-14, atol=1e-16)
    # the drift: far upstream the mean ux of the electrons is gamma v0
    far = up[0, :np2_row, :, 1] > ld + 2 + 5
    assert abs(up[2, :np2_row, :, 1][far].mean() - v0 / np.sqrt(1 - v0 * v0)) < 0.02


def test_shock_app_run():
    """the shock app loop (2d/proj/shock/app.f90:99-172) with the box fully expanded: injection each
    step, the particle and moment files at their cadence, the final restart"""
    nx, ny, n0, steps = 32, 8, 4, 6
    par = {"num_process": 1, "n_ppc": n0, "n_x": nx, "n_x_ini": nx - 1, "n_y": ny, "u_inject": 0.3,
           "mass_ratio": 25.0, "sigma_e": 0.1, "omega_pe": 0.1, "v_the": 0.05, "v_thi": 0.01, "theta_bn": 60.0,
           "phi_bn": 90.0, "l_damp_ini": 6.0, "seed": 3}
    cfg = {"config": {"verbose": 0, "datadir": ".", "max_elapsed": 1.0e6, "max_it": steps, "intvl_ptcl": steps,
                      "intvl_mom": 3, "intvl_orb": 100, "intvl_expand": 1}, "parameter": par}
    print("shock app run", flush=True)
    with tempfile.TemporaryDirectory() as d:
        path = _run_app(cfg, d)
        names = set(os.listdir(d))
        meta = json.load(open(os.path.join(d, "%07d_ptcl.json" % steps)))
        cfg2 = json.load(open(path))
    for n in ("%07d_mom.json" % 3, "%07d_mom.json" % 6, "%07d_ptcl.raw" % 6, "%07d_restart.raw" % (steps + 1),
              "init_param.json"):
        assert n in names, n
    assert cfg2["config"]["restart_file"] == "%07d_restart" % (steps + 1)
    v0 = 0.3 / np.sqrt(1 + 0.09)
    pflux = n0 * v0 * 1.0 * 1.0 * ny
    n_ion = meta["dataset"]["up01"]["shape"][1]
    n_init = n0 * (nx - 2) * ny
    assert n_init + steps * int(pflux) <= n_ion <= n_init + steps * (int(pflux) + 1)


def test_shock_tracers_in_the_initial_state_and_orb_files():
    """set_particle_ids (2d/proj/shock/app.f90:500-524): the ids of particles
    starting at 450 <= x <= 500 are positive, so the orbit files carry them
    (io__orb writes positive ids only)"""
    nx, ny, n0, steps = 520, 4, 1, 2
    par = {"num_process": 1, "n_ppc": n0, "n_x": nx, "n_x_ini": nx - 1, "n_y": ny, "u_inject": 0.3,
           "mass_ratio": 25.0, "sigma_e": 0.1, "omega_pe": 0.1, "v_the": 0.05, "v_thi": 0.01, "theta_bn": 60.0,
           "phi_bn": 90.0, "l_damp_ini": 6.0, "seed": 5}
    cfg = {"config": {"verbose": 0, "datadir": ".", "max_elapsed": 1.0e6, "max_it": steps, "intvl_ptcl": 100,
                      "intvl_mom": 100, "intvl_orb": 1, "intvl_expand": 1}, "parameter": par}
    np2_row = n0 * (nx - 2)
    ii = np.arange(1, np2_row + 1)
    x = 2 + (nx - 1) * (ii - 0.5) / np2_row
    want = int(((x >= 450.0) & (x <= 500.0)).sum()) * ny
    assert want > 0
    print("shock tracers", flush=True)
    with tempfile.TemporaryDirectory() as d:
        _run_app(cfg, d)
        for it in (1, 2):
            meta = json.load(open(os.path.join(d, "%07d_orb.json" % it)))
            for sp in ("up01", "up02"):
                assert meta["dataset"][sp]["shape"][1] == want, (it, sp, meta["dataset"][sp]["shape"])


def test_shock_app_expanding_box():
    """relocate in the shock app's loop (2d/proj/shock/app.f90:120-125, :615-692): the box starts
    with nxe = nxgs + n_x_ini and grows by one cell every intvl_expand steps; the restart records
    the grown edge; the particles stay left of it; the fields beyond nxe+2 keep their initial
    values; each relocation adds n0 particles per row per species"""
    nx, ny, n0, steps, ini, every = 40, 8, 4, 6, 20, 2
    par = {"num_process": 1, "n_ppc": n0, "n_x": nx, "n_x_ini": ini, "n_y": ny, "u_inject": 0.3,
           "mass_ratio": 25.0, "sigma_e": 0.1, "omega_pe": 0.1, "v_the": 0.05, "v_thi": 0.01, "theta_bn": 60.0,
           "phi_bn": 90.0, "l_damp_ini": 6.0, "seed": 3}
    np_cap = n0 * nx * 5
    print("shock app expanding box", flush=True)
    out = {}
    for max_it in (0, steps):
        cfg = {"config": {"verbose": 0, "datadir": ".", "max_elapsed": 1.0e6, "max_it": max_it,
                          "intvl_ptcl": 1000, "intvl_mom": 1000, "intvl_orb": 1000, "intvl_expand": every},
               "parameter": par}
        with tempfile.TemporaryDirectory() as d:
            _run_app(cfg, d)
            out[max_it] = _read_restart(d, "%07d_restart" % (max_it + 1), np_cap, ny, nx)
    meta0, up0, np20, uf0 = out[0]
    meta, up, np2, uf = out[steps]
    assert meta0["attribute"]["nxe"]["data"] == 2 + ini
    nxe = meta["attribute"]["nxe"]["data"]
    assert nxe == 2 + ini + steps // every
    assert np.all(np20 == n0 * (ini - 1))
    v0 = 0.3 / np.sqrt(1 + 0.09)
    pflux = n0 * v0 * 1.0 * 1.0 * ny
    for s in range(2):
        tot = int(np2[:, s].sum())
        base = n0 * (ini - 1) * ny + (steps // every) * n0 * ny
        assert base + steps * int(pflux) <= tot <= base + steps * (int(pflux) + 1)
        for j in range(ny):
            x = up[0, :np2[j, s], j, s]
            assert np.all(x >= 3.0) and np.all(x < nxe + 1.0)
    i_beyond = nxe + 3 - 0  # plane column of node nxe+3 (columns start at nxgs-2 = 0)
    assert np.array_equal(uf[:, i_beyond:, :].view(np.uint64), uf0[:, i_beyond:, :].view(np.uint64))
    assert not np.array_equal(uf[:, :i_beyond, :], uf0[:, :i_beyond, :])
