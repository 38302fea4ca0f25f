"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, never by the product package.

Arrays are numpy arrays in the reference's own Fortran layout (order='F'):
    up(ndim, np, nys:nye, nsp), uf(6, nxgs-2:nxge+2, nys-2:nye+2),
    cumcnt(nxgs:nxge+1, nys:nye, nsp), np2(nys:nye, nsp).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER2D = os.path.join(HERE, "_ref", "ref_golden2d")
REF_DRIVER3D = os.path.join(HERE, "_ref", "ref_golden3d")

MAXSP = 4
BC_PERIODIC = 0
BC_WALLX = 1
BC_SHOCK = 2


class Params2D(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("np", C.c_int), ("nsp", C.c_int),
        ("nxgs", C.c_int), ("nxge", C.c_int), ("nygs", C.c_int), ("nyge", C.c_int),
        ("nys", C.c_int), ("nye", C.c_int),
        ("delx", C.c_double), ("delt", C.c_double), ("c", C.c_double), ("gfac", C.c_double),
        ("q", C.c_double * MAXSP), ("r", C.c_double * MAXSP),
        ("bc", C.c_int), ("u0", C.c_double), ("nxe", C.c_int),
    ]

    @property
    def nx(self):
        return self.nxge - self.nxgs + 1

    @property
    def ny(self):
        return self.nye - self.nys + 1


class FState2D(C.Structure):
    _fields_ = [("df", C.POINTER(C.c_double)), ("gkl", C.POINTER(C.c_double)),
                ("uj", C.POINTER(C.c_double)), ("cg_iters", C.c_int * 3)]


class Params3D(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("np", C.c_int), ("nsp", C.c_int),
        ("nxgs", C.c_int), ("nxge", C.c_int), ("nygs", C.c_int), ("nyge", C.c_int),
        ("nzgs", C.c_int), ("nzge", C.c_int),
        ("nys", C.c_int), ("nye", C.c_int), ("nzs", C.c_int), ("nze", C.c_int),
        ("delx", C.c_double), ("delt", C.c_double), ("c", C.c_double), ("gfac", C.c_double),
        ("q", C.c_double * MAXSP), ("r", C.c_double * MAXSP),
    ]

    @property
    def nx(self):
        return self.nxge - self.nxgs + 1

    @property
    def ny(self):
        return self.nye - self.nys + 1

    @property
    def nz(self):
        return self.nze - self.nzs + 1


FState3D = FState2D


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def have_ref():
    return os.path.exists(REF_DRIVER2D)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
        P = C.POINTER(Params2D)
        S = C.POINTER(FState2D)
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_cg_err.argtypes = [C.c_double]
        L.or_cg_err.restype = C.c_double
        L.or_last_error.restype = C.c_char_p
        L.or2d_fstate_init.argtypes = [P, S]
        L.or2d_fstate_free.argtypes = [S]
        L.or2d_field_consts.argtypes = [P, np.ctypeslib.ndpointer(np.float64)]
        L.or2d_particle_solv.argtypes = [P, dp, dp, dp, ip]
        L.or2d_particle_solv_vay.argtypes = [P, dp, dp, dp, ip]
        L.or2d_bc_injection.argtypes = [P, dp, ip]
        L.or2d_ele_cur.argtypes = [P, dp, dp, dp, ip]
        L.or2d_bc_curre.argtypes = [P, dp]
        L.or2d_bc_dfield.argtypes = [P, dp]
        L.or2d_bc_phi.argtypes = [P, dp, C.c_int]
        L.or2d_field_fdtd_i.argtypes = [P, S, dp, dp, dp, ip]
        L.or2d_cgm.argtypes = [P, S]
        L.or2d_field_from_current.argtypes = [P, S, dp]
        L.or2d_bc_particle_x.argtypes = [P, dp, ip]
        L.or2d_bc_particle_y.argtypes = [P, dp, ip]
        L.or2d_sort_bucket.argtypes = [P, dp, dp, ip, ip]
        L.or2d_step.argtypes = [P, S, dp, dp, dp, ip, ip, C.c_int]
        L.or2d_energy.argtypes = [P, dp, dp, ip, np.ctypeslib.ndpointer(np.float64)]
        L.or2d_gauss.argtypes = [P, dp, dp, dp, ip]
        L.or2d_mom_accl.argtypes = [P, dp, dp, dp, ip]
        L.or2d_mom_nvt.argtypes = [P, dp, dp, ip]
        L.or2d_bc_mom.argtypes = [P, dp]
        P3 = C.POINTER(Params3D)
        L.or3d_fstate_init.argtypes = [P3, S]
        L.or3d_fstate_free.argtypes = [S]
        L.or3d_field_consts.argtypes = [P3, np.ctypeslib.ndpointer(np.float64)]
        L.or3d_particle_solv.argtypes = [P3, dp, dp, dp, ip]
        L.or3d_particle_solv_vay.argtypes = [P3, dp, dp, dp, ip]
        L.or3d_ele_cur.argtypes = [P3, dp, dp, dp, ip]
        L.or3d_bc_curre.argtypes = [P3, dp]
        L.or3d_bc_dfield.argtypes = [P3, dp]
        L.or3d_bc_phi.argtypes = [P3, dp, C.c_int]
        L.or3d_field_fdtd_i.argtypes = [P3, S, dp, dp, dp, ip]
        L.or3d_cgm.argtypes = [P3, S]
        L.or3d_field_from_current.argtypes = [P3, S, dp]
        L.or3d_bc_particle_x.argtypes = [P3, dp, ip]
        L.or3d_bc_particle_yz.argtypes = [P3, dp, ip]
        L.or3d_sort_bucket.argtypes = [P3, dp, dp, ip, ip]
        L.or3d_step.argtypes = [P3, S, dp, dp, dp, ip, ip, C.c_int]
        L.or3d_energy.argtypes = [P3, dp, dp, ip, np.ctypeslib.ndpointer(np.float64)]
        L.or3d_gauss.argtypes = [P3, dp, dp, dp, ip]
        L.or3d_mom_accl.argtypes = [P3, dp, dp, dp, ip]
        L.or3d_mom_nvt.argtypes = [P3, dp, dp, ip]
        L.or3d_bc_mom.argtypes = [P3, dp]
        for f in ("or_add_rd", "or_mul_rd", "or_div_rd"):
            getattr(L, f).argtypes = [C.c_double, C.c_double]
            getattr(L, f).restype = C.c_double
        _lib = L
    return _lib


def set_cg_err(err):
    """cgm's relative stopping tolerance for every later solve (the
    reference's 1e-6 by default); returns the previous value"""
    L = lib()
    old = L.or_cg_err()
    L.or_set_cg_err(err)
    return old


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


def make_params(nx, ny, nsp=2, ndim=6, np_cap=None, delx=1.0, delt=0.5, c=1.0, gfac=0.501,
                q=(0.0, 0.0), r=(1.0, 1.0), bc=BC_PERIODIC, n_ppc=None, u0=0.0):
    p = Params2D()
    p.ndim, p.nsp = ndim, nsp
    p.nxgs, p.nxge = 2, 2 + nx - 1
    p.nygs, p.nyge = 2, 2 + ny - 1
    p.nys, p.nye = p.nygs, p.nyge
    if np_cap is None:
        np_cap = 5 * (n_ppc or 16) * nx
    p.np = np_cap
    p.delx, p.delt, p.c, p.gfac = delx, delt, c, gfac
    for k in range(nsp):
        p.q[k] = q[k]
        p.r[k] = r[k]
    p.bc = bc
    p.u0 = u0
    return p


def alloc_state(p):
    """Zeroed reference-layout arrays: up, gp, uf, cumcnt, np2."""
    up = np.zeros((p.ndim, p.np, p.ny, p.nsp), dtype=np.float64, order="F")
    gp = np.zeros_like(up, order="F")
    uf = np.zeros((6, p.nx + 4, p.ny + 4), dtype=np.float64, order="F")
    cumcnt = np.zeros((p.nx + 1, p.ny, p.nsp), dtype=np.int32, order="F")
    np2 = np.zeros((p.ny, p.nsp), dtype=np.int32, order="F")
    return up, gp, uf, cumcnt, np2


class Oracle2D:
    """The oracle's persistent field state (df, gkl, uj) plus entry points."""

    def __init__(self, p, threads=1):
        self.p = p
        self.L = lib()
        self.L.or_set_threads(threads)
        self.st = FState2D()
        _check(self.L.or2d_fstate_init(C.byref(p), C.byref(self.st)))
        nx, ny = p.nx, p.ny
        self.df = np.ctypeslib.as_array(self.st.df, shape=(6 * (nx + 4) * (ny + 4),)).reshape(
            (6, nx + 4, ny + 4), order="F")
        self.uj = np.ctypeslib.as_array(self.st.uj, shape=(3 * (nx + 4) * (ny + 4),)).reshape(
            (3, nx + 4, ny + 4), order="F")
        self.gkl = np.ctypeslib.as_array(self.st.gkl, shape=(3 * nx * ny,)).reshape((3, nx, ny), order="F")

    def __del__(self):
        try:
            self.L.or2d_fstate_free(C.byref(self.st))
        except Exception:
            pass

    def particle_solv(self, gp, up, uf, cumcnt):
        self.L.or2d_particle_solv(C.byref(self.p), gp, up, uf, cumcnt)

    def particle_solv_vay(self, gp, up, uf, cumcnt):
        self.L.or2d_particle_solv_vay(C.byref(self.p), gp, up, uf, cumcnt)

    def ele_cur(self, uj, up, gp, cumcnt):
        self.L.or2d_ele_cur(C.byref(self.p), uj, up, gp, cumcnt)

    def bc_curre(self, uj):
        self.L.or2d_bc_curre(C.byref(self.p), uj)

    def bc_dfield(self, df):
        self.L.or2d_bc_dfield(C.byref(self.p), df)

    def bc_phi(self, phi, l):
        self.L.or2d_bc_phi(C.byref(self.p), phi, l)

    def field_fdtd_i(self, uf, up, gp, cumcnt):
        _check(self.L.or2d_field_fdtd_i(C.byref(self.p), C.byref(self.st), uf, up, gp, cumcnt))
        return tuple(self.st.cg_iters)

    def cgm(self):
        _check(self.L.or2d_cgm(C.byref(self.p), C.byref(self.st)))
        return tuple(self.st.cg_iters)

    def field_from_current(self, uf):
        """field__fdtd_i after ele_cur and the current fold, from self.uj"""
        _check(self.L.or2d_field_from_current(C.byref(self.p), C.byref(self.st), uf))
        return tuple(self.st.cg_iters)

    def bc_particle_x(self, up, np2):
        self.L.or2d_bc_particle_x(C.byref(self.p), up, np2)

    def bc_injection(self, up, np2):
        self.L.or2d_bc_injection(C.byref(self.p), up, np2)

    def bc_particle_y(self, up, np2):
        _check(self.L.or2d_bc_particle_y(C.byref(self.p), up, np2))

    def sort_bucket(self, gp, up, cumcnt, np2):
        self.L.or2d_sort_bucket(C.byref(self.p), gp, up, cumcnt, np2)

    def step(self, up, gp, uf, cumcnt, np2, nsteps=1):
        _check(self.L.or2d_step(C.byref(self.p), C.byref(self.st), up, gp, uf, cumcnt, np2, nsteps))

    def energy(self, up, uf, np2):
        out = np.zeros(3)
        self.L.or2d_energy(C.byref(self.p), up, uf, np2, out)
        return out

    def gauss(self, uf, up, np2):
        G = np.zeros((self.p.nx, self.p.ny), order="F")
        self.L.or2d_gauss(C.byref(self.p), G, uf, up, np2)
        return G

    def mom_accl(self, gp, up, uf, cumcnt):
        self.L.or2d_mom_accl(C.byref(self.p), gp, up, uf, cumcnt)

    def mom_nvt(self, up, np2):
        """mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp) of up's first np2 slots."""
        mom = np.zeros((7, self.p.nx + 2, self.p.ny + 2, self.p.nsp), order="F")
        self.L.or2d_mom_nvt(C.byref(self.p), mom, up, np2)
        return mom

    def bc_mom(self, mom):
        self.L.or2d_bc_mom(C.byref(self.p), mom)


def run_ref2d(case_dir, mode, p, up, np2=None, uf=None, cumcnt=None):
    """Run the reference's own particle__solv (mode 1), sort__bucket (mode 2),
    mom_calc__accl + mom_calc__nvt (mode 3: returns gp, mom) or
    particle__solv_vay (mode 4) through
    oracle/_ref/ref_golden2d on raw files in case_dir."""
    os.makedirs(case_dir, exist_ok=True)
    with open(os.path.join(case_dir, "params.txt"), "w") as f:
        f.write(f"{mode} {p.ndim} {p.np} {p.nsp} {p.nxgs} {p.nxge} {p.nygs} {p.nyge} {p.nys} {p.nye}\n")
        f.write(f"{p.delx!r} {p.delt!r} {p.c!r}\n")
        f.write(" ".join(repr(float(p.q[k])) for k in range(p.nsp)) + "\n")
        f.write(" ".join(repr(float(p.r[k])) for k in range(p.nsp)) + "\n")
    up.ravel(order="F").tofile(os.path.join(case_dir, "up.bin"))
    if mode in (1, 3, 4):
        uf.ravel(order="F").tofile(os.path.join(case_dir, "uf.bin"))
        cumcnt.ravel(order="F").tofile(os.path.join(case_dir, "cumcnt.bin"))
    if mode in (2, 3):
        np2.ravel(order="F").tofile(os.path.join(case_dir, "np2.bin"))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([REF_DRIVER2D, case_dir], check=True, env=env)
    gp = np.fromfile(os.path.join(case_dir, "gp.bin"), dtype=np.float64).reshape(up.shape, order="F")
    if mode in (1, 4):
        return gp
    if mode == 3:
        mom = np.fromfile(os.path.join(case_dir, "mom.bin"), dtype=np.float64).reshape(
            (7, p.nx + 2, p.ny + 2, p.nsp), order="F")
        return gp, mom
    cc = np.fromfile(os.path.join(case_dir, "cumcnt.bin"), dtype=np.int32).reshape(
        (p.nx + 1, p.ny, p.nsp), order="F")
    return gp, cc


# ---------------------------------------------------------------------------
# 3D: up(7, np, nys:nye, nzs:nze, nsp), uf(6, nxgs-2:nxge+2, nys-2:nye+2,
# nzs-2:nze+2), cumcnt(nxgs:nxge+1, nys:nye, nzs:nze, nsp), np2(nys:nye, nzs:nze, nsp)
# ---------------------------------------------------------------------------
def make_params3d(nx, ny, nz, nsp=2, ndim=7, np_cap=None, delx=1.0, delt=0.5, c=1.0, gfac=0.501,
                  q=(0.0, 0.0), r=(1.0, 1.0), n_ppc=None):
    p = Params3D()
    p.ndim, p.nsp = ndim, nsp
    p.nxgs, p.nxge = 2, 2 + nx - 1
    p.nygs, p.nyge = 2, 2 + ny - 1
    p.nzgs, p.nzge = 2, 2 + nz - 1
    p.nys, p.nye = p.nygs, p.nyge
    p.nzs, p.nze = p.nzgs, p.nzge
    if np_cap is None:
        np_cap = 3 * (n_ppc or 16) * nx
    p.np = np_cap
    p.delx, p.delt, p.c, p.gfac = delx, delt, c, gfac
    for k in range(nsp):
        p.q[k] = q[k]
        p.r[k] = r[k]
    return p


def alloc_state3d(p):
    """Zeroed reference-layout 3D arrays: up, gp, uf, cumcnt, np2."""
    up = np.zeros((p.ndim, p.np, p.ny, p.nz, p.nsp), dtype=np.float64, order="F")
    gp = np.zeros_like(up, order="F")
    uf = np.zeros((6, p.nx + 4, p.ny + 4, p.nz + 4), dtype=np.float64, order="F")
    cumcnt = np.zeros((p.nx + 1, p.ny, p.nz, p.nsp), dtype=np.int32, order="F")
    np2 = np.zeros((p.ny, p.nz, p.nsp), dtype=np.int32, order="F")
    return up, gp, uf, cumcnt, np2


class Oracle3D:
    """3D twin of Oracle2D (the 3d/common modules at one rank, periodic)."""

    def __init__(self, p, threads=1):
        self.p = p
        self.L = lib()
        self.L.or_set_threads(threads)
        self.st = FState3D()
        _check(self.L.or3d_fstate_init(C.byref(p), C.byref(self.st)))
        nx, ny, nz = p.nx, p.ny, p.nz
        n4 = (nx + 4) * (ny + 4) * (nz + 4)
        self.df = np.ctypeslib.as_array(self.st.df, shape=(6 * n4,)).reshape((6, nx + 4, ny + 4, nz + 4), order="F")
        self.uj = np.ctypeslib.as_array(self.st.uj, shape=(3 * n4,)).reshape((3, nx + 4, ny + 4, nz + 4), order="F")
        self.gkl = np.ctypeslib.as_array(self.st.gkl, shape=(3 * nx * ny * nz,)).reshape((3, nx, ny, nz), order="F")

    def __del__(self):
        try:
            self.L.or3d_fstate_free(C.byref(self.st))
        except Exception:
            pass

    def field_consts(self):
        f = np.zeros(5)
        self.L.or3d_field_consts(C.byref(self.p), f)
        return f

    def particle_solv(self, gp, up, uf, cumcnt):
        self.L.or3d_particle_solv(C.byref(self.p), gp, up, uf, cumcnt)

    def particle_solv_vay(self, gp, up, uf, cumcnt):
        self.L.or3d_particle_solv_vay(C.byref(self.p), gp, up, uf, cumcnt)

    def ele_cur(self, uj, up, gp, cumcnt):
        self.L.or3d_ele_cur(C.byref(self.p), uj, up, gp, cumcnt)

    def bc_curre(self, uj):
        self.L.or3d_bc_curre(C.byref(self.p), uj)

    def bc_dfield(self, df):
        self.L.or3d_bc_dfield(C.byref(self.p), df)

    def bc_phi(self, phi, l):
        self.L.or3d_bc_phi(C.byref(self.p), phi, l)

    def field_fdtd_i(self, uf, up, gp, cumcnt):
        _check(self.L.or3d_field_fdtd_i(C.byref(self.p), C.byref(self.st), uf, up, gp, cumcnt))
        return tuple(self.st.cg_iters)

    def cgm(self):
        _check(self.L.or3d_cgm(C.byref(self.p), C.byref(self.st)))
        return tuple(self.st.cg_iters)

    def field_from_current(self, uf):
        """field__fdtd_i after ele_cur and the current fold, from self.uj"""
        _check(self.L.or3d_field_from_current(C.byref(self.p), C.byref(self.st), uf))
        return tuple(self.st.cg_iters)

    def bc_particle_x(self, up, np2):
        self.L.or3d_bc_particle_x(C.byref(self.p), up, np2)

    def bc_particle_yz(self, up, np2):
        _check(self.L.or3d_bc_particle_yz(C.byref(self.p), up, np2))

    def sort_bucket(self, gp, up, cumcnt, np2):
        self.L.or3d_sort_bucket(C.byref(self.p), gp, up, cumcnt, np2)

    def step(self, up, gp, uf, cumcnt, np2, nsteps=1):
        _check(self.L.or3d_step(C.byref(self.p), C.byref(self.st), up, gp, uf, cumcnt, np2, nsteps))

    def energy(self, up, uf, np2):
        out = np.zeros(3)
        self.L.or3d_energy(C.byref(self.p), up, uf, np2, out)
        return out

    def gauss(self, uf, up, np2):
        G = np.zeros((self.p.nx, self.p.ny, self.p.nz), order="F")
        self.L.or3d_gauss(C.byref(self.p), G, uf, up, np2)
        return G

    def mom_accl(self, gp, up, uf, cumcnt):
        self.L.or3d_mom_accl(C.byref(self.p), gp, up, uf, cumcnt)

    def mom_nvt(self, up, np2):
        """mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp)."""
        p = self.p
        mom = np.zeros((7, p.nx + 2, p.ny + 2, p.nz + 2, p.nsp), order="F")
        self.L.or3d_mom_nvt(C.byref(p), mom, up, np2)
        return mom

    def bc_mom(self, mom):
        self.L.or3d_bc_mom(C.byref(self.p), mom)


def have_ref3d():
    return os.path.exists(REF_DRIVER3D)


def run_ref3d(case_dir, mode, p, up, np2=None, uf=None, cumcnt=None):
    """Run the reference's own 3D particle__solv (mode 1), sort__bucket (mode 2),
    mom_calc__accl + mom_calc__nvt (mode 3: returns gp, mom) or particle__solv_vay
    (mode 4) through
    oracle/_ref/ref_golden3d on raw files in case_dir."""
    os.makedirs(case_dir, exist_ok=True)
    with open(os.path.join(case_dir, "params.txt"), "w") as f:
        f.write(f"{mode} {p.ndim} {p.np} {p.nsp} {p.nxgs} {p.nxge} {p.nygs} {p.nyge} {p.nzgs} {p.nzge} "
                f"{p.nys} {p.nye} {p.nzs} {p.nze}\n")
        f.write(f"{p.delx!r} {p.delt!r} {p.c!r}\n")
        f.write(" ".join(repr(float(p.q[k])) for k in range(p.nsp)) + "\n")
        f.write(" ".join(repr(float(p.r[k])) for k in range(p.nsp)) + "\n")
    up.ravel(order="F").tofile(os.path.join(case_dir, "up.bin"))
    if mode in (1, 3, 4):
        uf.ravel(order="F").tofile(os.path.join(case_dir, "uf.bin"))
        cumcnt.ravel(order="F").tofile(os.path.join(case_dir, "cumcnt.bin"))
    if mode in (2, 3):
        np2.ravel(order="F").tofile(os.path.join(case_dir, "np2.bin"))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([REF_DRIVER3D, case_dir], check=True, env=env)
    gp = np.fromfile(os.path.join(case_dir, "gp.bin"), dtype=np.float64).reshape(up.shape, order="F")
    if mode in (1, 4):
        return gp
    if mode == 3:
        mom = np.fromfile(os.path.join(case_dir, "mom.bin"), dtype=np.float64).reshape(
            (7, p.nx + 2, p.ny + 2, p.nz + 2, p.nsp), order="F")
        return gp, mom
    cc = np.fromfile(os.path.join(case_dir, "cumcnt.bin"), dtype=np.int32).reshape(
        (p.nx + 1, p.ny, p.nz, p.nsp), order="F")
    return gp, cc
