"""Host-side mirror of the reference's 3D module entry points over the C-ABI
(include/wpic3d.h), periodic in x, y and z at one rank.

`Pic3D` owns one device context; its methods carry the reference's names and
argument meaning (3d/common/particle.f90, field.f90, sort.f90,
boundary_periodic.f90).  Arrays cross the boundary only through the
upload_*/download_* methods, in the reference's own Fortran layouts:

    up(7, np, nys:nye, nzs:nze, nsp), uf(6, nxgs-2:nxge+2, nys-2:nye+2, nzs-2:nze+2),
    cumcnt(nxgs:nxge+1, nys:nye, nzs:nze, nsp), np2(nys:nye, nzs:nze, nsp)
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import MODE_FAST, MODE_PARITY, check  # noqa: F401
from .pic2d import weibel_charges


class Pic3D:
    def __init__(self, nx, ny, nz, q, r, *, nsp=2, delx=1.0, delt=0.5, c=1.0, gfac=0.501, mode=MODE_FAST,
                 capacity=None, nxgs=2, nygs=2, nzgs=2, n_ppc_hint=16):
        L = _lib.lib()
        p = _lib.Params3D()
        p.ndim, p.nsp = 7, nsp
        p.nxgs, p.nxge = nxgs, nxgs + nx - 1
        p.nygs, p.nyge = nygs, nygs + ny - 1
        p.nzgs, p.nzge = nzgs, nzgs + nz - 1
        p.delx, p.delt, p.c, p.gfac = delx, delt, c, gfac
        for k in range(nsp):
            p.q[k] = q[k]
            p.r[k] = r[k]
        p.mode = mode
        p.capacity = int(capacity if capacity is not None else n_ppc_hint * nx * ny * nz)
        self.p = p
        self.nx, self.ny, self.nz, self.nsp = nx, ny, nz, nsp
        self._L = L
        self._h = C.c_void_p()
        check(L.wpic_create3d(C.byref(self._h), C.byref(p)))

    def close(self):
        if self._h:
            self._L.wpic_destroy3d(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, mode):
        check(self._L.wpic_set_mode3d(self._h, mode))

    def sync(self):
        check(self._L.wpic_sync3d(self._h))

    # -- reference-layout transfers ---------------------------------------------
    def _fshape(self, ncomp):
        return (ncomp, self.nx + 4, self.ny + 4, self.nz + 4)

    def upload_up(self, up, np2, cumcnt):
        up = np.asfortranarray(up, dtype=np.float64)
        check(self._L.wpic_upload_up3d(self._h, up, np.asfortranarray(np2, dtype=np.int32),
                                       np.asfortranarray(cumcnt, dtype=np.int32), up.shape[1]))

    def download_up(self, np_cap):
        up = np.zeros((7, np_cap, self.ny, self.nz, self.nsp), order="F")
        cumcnt = np.zeros((self.nx + 1, self.ny, self.nz, self.nsp), dtype=np.int32, order="F")
        np2 = np.zeros((self.ny, self.nz, self.nsp), dtype=np.int32, order="F")
        check(self._L.wpic_download_up3d(self._h, up, np2, cumcnt, np_cap))
        return up, np2, cumcnt

    def upload_gp(self, gp, np2):
        gp = np.asfortranarray(gp, dtype=np.float64)
        check(self._L.wpic_upload_gp3d(self._h, gp, np.asfortranarray(np2, dtype=np.int32), gp.shape[1]))

    def download_gp(self, np_cap):
        gp = np.zeros((7, np_cap, self.ny, self.nz, self.nsp), order="F")
        np2 = np.zeros((self.ny, self.nz, self.nsp), dtype=np.int32, order="F")
        check(self._L.wpic_download_gp3d(self._h, gp, np2, np_cap))
        return gp, np2

    def upload_fields(self, uf):
        check(self._L.wpic_upload_fields3d(self._h, np.asfortranarray(uf, dtype=np.float64)))

    def download_fields(self):
        uf = np.zeros(self._fshape(6), order="F")
        check(self._L.wpic_download_fields3d(self._h, uf))
        return uf

    def download_current(self):
        uj = np.zeros(self._fshape(3), order="F")
        check(self._L.wpic_download_current3d(self._h, uj))
        return uj

    def download_dfield(self):
        df = np.zeros(self._fshape(6), order="F")
        check(self._L.wpic_download_dfield3d(self._h, df))
        return df

    def upload_dfield(self, df):
        check(self._L.wpic_upload_dfield3d(self._h, np.asfortranarray(df, dtype=np.float64)))

    # -- reference entry points ---------------------------------------------------
    def particle_solv_vay(self):
        """particle__solv_vay(gp, up, uf, cumcnt, nxs, nxe): 3d/common/particle.f90:236"""
        check(self._L.wpic_particle_solv_vay3d(self._h))

    def set_pusher(self, pusher):
        """the kick of particle_solv and step: _lib.PUSHER_BORIS or _lib.PUSHER_VAY"""
        check(self._L.wpic_set_pusher3d(self._h, int(pusher)))

    def particle_solv(self):
        """particle__solv(gp, up, uf, cumcnt, nxs, nxe): 3d/common/particle.f90:52"""
        check(self._L.wpic_particle_solv3d(self._h))

    def field_fdtd_i(self):
        """field__fdtd_i(uf, up, gp, cumcnt, ...): 3d/common/field.f90:70; returns
        the CG iteration count of each component"""
        check(self._L.wpic_field_fdtd_i3d(self._h))
        return self.cg_iters()

    def bc_particle_x(self):
        """boundary_periodic__particle_x(gp, np2): 3d/common/boundary_periodic.f90:68"""
        check(self._L.wpic_bc_particle_x3d(self._h))

    def bc_particle_yz(self):
        """boundary_periodic__particle_yz(gp, np2): 3d/common/boundary_periodic.f90:104"""
        check(self._L.wpic_bc_particle_yz3d(self._h))

    def sort_bucket(self):
        """sort__bucket(up, gp, cumcnt, np2, nxs, nxe): 3d/common/sort.f90:40"""
        check(self._L.wpic_sort_bucket3d(self._h))

    def cg_iters(self):
        it = np.zeros(3, dtype=np.int32)
        check(self._L.wpic_cg_iters3d(self._h, it))
        return tuple(int(v) for v in it)

    def set_cg_tol(self, err):
        """the CG's relative stopping tolerance (1e-6 as in field.f90 by default)"""
        check(self._L.wpic_set_cg_tol3d(self._h, float(err)))

    def step(self, nsteps=1):
        """nsteps of the loop body of 3d/proj/weibel/app.f90:103-108"""
        check(self._L.wpic_step3d(self._h, int(nsteps)))

    # -- moments (3d/common/mom_calc.f90, boundary_periodic__mom) ----------------------
    def mom_accl(self):
        """mom_calc__accl(gp, up, uf, cumcnt, nxs, nxe): 3d/common/mom_calc.f90:49"""
        check(self._L.wpic_mom_accl3d(self._h))

    def mom_nvt(self):
        """mom_calc__nvt(mom, gp, np2): 3d/common/mom_calc.f90:219"""
        check(self._L.wpic_mom_nvt3d(self._h))

    def bc_mom(self):
        """boundary_periodic__mom(mom): 3d/common/boundary_periodic.f90:1102"""
        check(self._L.wpic_bc_mom3d(self._h))

    def moments(self):
        """accl, nvt and bc__mom in one pass (3d/proj/weibel/app.f90:122-124)"""
        check(self._L.wpic_moments3d(self._h))

    def download_mom(self):
        """mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp)"""
        mom = np.zeros((7, self.nx + 2, self.ny + 2, self.nz + 2, self.nsp), order="F")
        check(self._L.wpic_download_mom3d(self._h, mom))
        return mom

    def upload_mom(self, mom):
        check(self._L.wpic_upload_mom3d(self._h, np.asfortranarray(mom, dtype=np.float64)))

    # -- diagnostics and initial state ------------------------------------------------
    def energy(self):
        out = np.zeros(3)
        check(self._L.wpic_energy3d(self._h, out))
        return out

    def gauss(self):
        G = np.zeros((self.nx, self.ny, self.nz), order="F")
        check(self._L.wpic_gauss3d(self._h, G))
        return G

    def counts(self):
        n = np.zeros(_lib.MAXSP, dtype=np.int64)
        check(self._L.wpic_counts3d(self._h, n))
        return n[: self.nsp]

    def ic_beams(self, seed, n_ppc, udrift, uth_e, uth_i, b0=0.0):
        check(self._L.wpic_ic_beams3d(self._h, int(seed), int(n_ppc), float(udrift), float(uth_e),
                                      float(uth_i), float(b0)))

    def energy_species(self):
        """[kinetic of each species..., E^2/8pi, B^2/8pi] (3d/proj/weibel/app.f90:509-577)"""
        out = np.zeros(self.nsp + 2)
        check(self._L.wpic_energy_species3d(self._h, out))
        return out

    def ic_weibel(self, seed, n0, v_thi, v_the, t_ani, b0):
        """the reference 3D Weibel app's state (3d/proj/weibel/app.f90 set_initial_condition)"""
        check(self._L.wpic_ic_weibel3d(self._h, int(seed), int(n0), float(v_thi), float(v_the), float(t_ani),
                                       float(b0)))

    def set_timing(self, on=True):
        check(self._L.wpic_set_timing3d(self._h, 1 if on else 0))

    def timing(self, name):
        ms, cnt = C.c_double(), C.c_longlong()
        check(self._L.wpic_timing3d(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


def weibel_beams_3d(nx, ny, nz, n_ppc, *, seed=20261202, omega_pe=0.1, mass_ratio=25.0, udrift=0.5,
                    uth_e=0.1, uth_i=0.01, delt=0.5, mode=MODE_FAST, capacity=None):
    """The 3D Weibel beam workload (3d/proj/weibel with dt = 0.5): two species,
    periodic box, seeded co-located in-cell positions, electron beams at
    u_x = +/-udrift with thermal spread uth_e, cold ions at uth_i."""
    q, r = weibel_charges(n_ppc, omega_pe, mass_ratio)
    cap = capacity if capacity is not None else n_ppc * nx * ny * nz
    sim = Pic3D(nx, ny, nz, q, r, delt=delt, mode=mode, capacity=cap)
    sim.ic_beams(seed, n_ppc, udrift, uth_e, uth_i, 0.0)
    return sim
