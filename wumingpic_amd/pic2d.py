"""Host-side mirror of the reference's 2D module entry points over the C-ABI.

`Pic2D` owns one device context.  Its methods carry the reference's names and
argument meaning (2d/common/particle.f90, field.f90, sort.f90,
boundary_periodic.f90, 2d/proj/reconnection/boundary_reconnection.f90); the
particle and field arrays stay resident on the GPU between calls and cross
the boundary only through the upload_*/download_* methods, in the
reference's own Fortran layouts:

    up(ndim, np, nys:nye, nsp), uf(6, nxgs-2:nxge+2, nys-2:nye+2),
    cumcnt(nxgs:nxge+1, nys:nye, nsp), np2(nys:nye, nsp)

Errors raise WpicError with the reference's message where the reference
stops with one (e.g. "********** stop at cgm after ite_max **********").
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BC_PERIODIC, BC_WALLX, MODE_FAST, MODE_PARITY, check  # noqa: F401


def weibel_charges(n_ppc, omega_pe, mass_ratio, delx=1.0):
    """q, r of 2d/proj/weibel/app.f90:293-302 with sigma_e = 0: ions first."""
    pi = np.pi
    wpe = omega_pe
    wpi = wpe / np.sqrt(mass_ratio)
    r = [mass_ratio, 1.0]
    q = [+np.sqrt(r[0] / (4 * pi * n_ppc / delx ** 2)) * wpi,
         -np.sqrt(r[1] / (4 * pi * n_ppc / delx ** 2)) * wpe]
    return q, r


class Pic2D:
    def __init__(self, nx, ny, q, r, *, nsp=2, delx=1.0, delt=0.5, c=1.0, gfac=0.501,
                 bc=BC_PERIODIC, mode=MODE_FAST, capacity=None, nxgs=2, nygs=2, n_ppc_hint=64):
        L = _lib.lib()
        p = _lib.Params2D()
        p.ndim, p.nsp = 6, nsp
        p.nxgs, p.nxge = nxgs, nxgs + nx - 1
        p.nygs, p.nyge = nygs, nygs + ny - 1
        p.delx, p.delt, p.c, p.gfac = delx, delt, c, gfac
        for k in range(nsp):
            p.q[k] = q[k]
            p.r[k] = r[k]
        p.bc, p.mode = bc, mode
        p.capacity = int(capacity if capacity is not None else n_ppc_hint * nx * ny)
        self.p = p
        self.nx, self.ny, self.nsp = nx, ny, nsp
        self._L = L
        self._h = C.c_void_p()
        check(L.wpic_create2d(C.byref(self._h), C.byref(p)))

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if self._h:
            self._L.wpic_destroy(self._h)
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
        check(self._L.wpic_set_mode(self._h, mode))

    def sync(self):
        check(self._L.wpic_sync(self._h))

    # -- reference-layout transfers -------------------------------------------
    def _shapes(self, np_cap):
        return ((6, np_cap, self.ny, self.nsp), (self.nx + 1, self.ny, self.nsp), (self.ny, self.nsp))

    def upload_up(self, up, np2, cumcnt):
        up = np.asfortranarray(up, dtype=np.float64)
        check(self._L.wpic_upload_up2d(self._h, up, np.asfortranarray(np2, dtype=np.int32),
                                       np.asfortranarray(cumcnt, dtype=np.int32), up.shape[1]))

    def download_up(self, np_cap, up=None):
        su, sc, sn = self._shapes(np_cap)
        up = np.zeros(su, order="F") if up is None else up
        cumcnt = np.zeros(sc, dtype=np.int32, order="F")
        np2 = np.zeros(sn, dtype=np.int32, order="F")
        check(self._L.wpic_download_up2d(self._h, up, np2, cumcnt, np_cap))
        return up, np2, cumcnt

    def upload_gp(self, gp, np2):
        gp = np.asfortranarray(gp, dtype=np.float64)
        check(self._L.wpic_upload_gp2d(self._h, gp, np.asfortranarray(np2, dtype=np.int32), gp.shape[1]))

    def download_gp(self, np_cap, gp=None):
        su, _, sn = self._shapes(np_cap)
        gp = np.zeros(su, order="F") if gp is None else gp
        np2 = np.zeros(sn, dtype=np.int32, order="F")
        check(self._L.wpic_download_gp2d(self._h, gp, np2, np_cap))
        return gp, np2

    def upload_fields(self, uf):
        check(self._L.wpic_upload_fields2d(self._h, np.asfortranarray(uf, dtype=np.float64)))

    def download_fields(self):
        uf = np.zeros((6, self.nx + 4, self.ny + 4), order="F")
        check(self._L.wpic_download_fields2d(self._h, uf))
        return uf

    def download_current(self):
        uj = np.zeros((3, self.nx + 4, self.ny + 4), order="F")
        check(self._L.wpic_download_current2d(self._h, uj))
        return uj

    def download_dfield(self):
        df = np.zeros((6, self.nx + 4, self.ny + 4), order="F")
        check(self._L.wpic_download_dfield2d(self._h, df))
        return df

    def upload_dfield(self, df):
        check(self._L.wpic_upload_dfield2d(self._h, np.asfortranarray(df, dtype=np.float64)))

    # -- reference entry points -----------------------------------------------
    def particle_solv(self):
        """particle__solv(gp, up, uf, cumcnt, nxs, nxe): 2d/common/particle.f90:48"""
        check(self._L.wpic_particle_solv2d(self._h))

    def bc_injection(self, u0):
        """boundary_shock__injection(gp, np2, nxs, nxe, u0): 2d/proj/shock/boundary_shock.f90:255"""
        check(self._L.wpic_bc_injection2d(self._h, float(u0)))

    def inject(self, seed, it, nl, v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn):
        """the shock app's inject (2d/proj/shock/app.f90:697-856); nl[r] particles per species join row r"""
        nl = np.ascontiguousarray(nl, dtype=np.int32)
        if nl.shape != (self.ny,):
            raise ValueError(f"inject needs one count per row: nl of shape ({self.ny},), got {nl.shape}")
        check(self._L.wpic_inject2d(self._h, int(seed), int(it), nl,
                                    float(v0), float(v_thi), float(v_the), float(l_damp), float(b0),
                                    float(theta_bn), float(phi_bn)))

    def relocate(self, seed, it, n0, v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn):
        """the shock app's relocate (2d/proj/shock/app.f90:615-692): the box grows by one cell at
        its active edge nxe, each row of each species gains n0 particles in the new cell nxe-1"""
        check(self._L.wpic_relocate2d(self._h, int(seed), int(it), int(n0), float(v0), float(v_thi),
                                      float(v_the), float(l_damp), float(b0), float(theta_bn), float(phi_bn)))

    @property
    def nxe(self):
        """the active right edge nxe of the shock's box (nxge unless set or grown)"""
        v = C.c_int()
        check(self._L.wpic_get_nxe2d(self._h, C.byref(v)))
        return v.value

    @nxe.setter
    def nxe(self, nxe):
        check(self._L.wpic_set_nxe2d(self._h, int(nxe)))

    def ic_shock(self, seed, n0, v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn):
        """the shock app's initial state (2d/proj/shock/app.f90:405-470) over the active box"""
        check(self._L.wpic_ic_shock2d(self._h, int(seed), int(n0), float(v0), float(v_thi), float(v_the),
                                      float(l_damp), float(b0), float(theta_bn), float(phi_bn)))

    def set_injection(self, u0):
        """the u0 of the step's injection edge (WPIC_BC_SHOCK)"""
        check(self._L.wpic_set_injection2d(self._h, float(u0)))

    def particle_solv_vay(self):
        """particle__solv_vay(gp, up, uf, cumcnt, nxs, nxe): 2d/common/particle.f90:182"""
        check(self._L.wpic_particle_solv_vay2d(self._h))

    def set_pusher(self, pusher):
        """the kick of particle_solv and step: _lib.PUSHER_BORIS or _lib.PUSHER_VAY"""
        check(self._L.wpic_set_pusher2d(self._h, int(pusher)))

    def field_fdtd_i(self):
        """field__fdtd_i(uf, up, gp, cumcnt, ...): 2d/common/field.f90:66.
        Returns the CG iteration count of each component."""
        check(self._L.wpic_field_fdtd_i2d(self._h))
        return self.cg_iters()

    def bc_particle_x(self):
        check(self._L.wpic_bc_particle_x2d(self._h))

    def bc_particle_y(self):
        check(self._L.wpic_bc_particle_y2d(self._h))

    def sort_bucket(self):
        """sort__bucket(up, gp, cumcnt, np2, nxs, nxe): 2d/common/sort.f90:36"""
        check(self._L.wpic_sort_bucket2d(self._h))

    def cg_iters(self):
        it = np.zeros(3, dtype=np.int32)
        check(self._L.wpic_cg_iters2d(self._h, it))
        return tuple(int(v) for v in it)

    def step(self, nsteps=1):
        check(self._L.wpic_step2d(self._h, int(nsteps)))

    # -- moments (2d/common/mom_calc.f90, boundary_*__mom) ---------------------
    def mom_accl(self):
        """mom_calc__accl(gp, up, uf, cumcnt, nxs, nxe): mom_calc.f90:48"""
        check(self._L.wpic_mom_accl2d(self._h))

    def mom_nvt(self):
        """mom_calc__nvt(mom, gp, np2): mom_calc.f90:167"""
        check(self._L.wpic_mom_nvt2d(self._h))

    def bc_mom(self):
        """boundary_periodic__mom / boundary_reconnection__mom by the boundary kind"""
        check(self._L.wpic_bc_mom2d(self._h))

    def moments(self):
        """accl, nvt and bc__mom in one pass (2d/proj/weibel/app.f90:121-123)"""
        check(self._L.wpic_moments2d(self._h))

    def download_mom(self):
        """mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp)"""
        mom = np.zeros((7, self.nx + 2, self.ny + 2, self.nsp), order="F")
        check(self._L.wpic_download_mom2d(self._h, mom))
        return mom

    def upload_mom(self, mom):
        check(self._L.wpic_upload_mom2d(self._h, np.asfortranarray(mom, dtype=np.float64)))

    # -- diagnostics and initial state ------------------------------------------
    def energy(self):
        out = np.zeros(3)
        check(self._L.wpic_energy2d(self._h, out))
        return out

    def gauss(self):
        G = np.zeros((self.nx, self.ny), order="F")
        check(self._L.wpic_gauss2d(self._h, G))
        return G

    def counts(self):
        n = np.zeros(_lib.MAXSP, dtype=np.int64)
        check(self._L.wpic_counts2d(self._h, n))
        return n[: self.nsp]

    def ic_beams(self, seed, n_ppc, udrift, uth_e, uth_i, b0=0.0):
        check(self._L.wpic_ic_beams2d(self._h, int(seed), int(n_ppc), float(udrift), float(uth_e),
                                      float(uth_i), float(b0)))

    # -- adjoint (include/wpic_adj.h) ---------------------------------------------
    def ic_weibel(self, seed, n0, v_thi, v_the, t_ani, b0):
        """the reference Weibel app's state (2d/proj/weibel/app.f90:366-438)"""
        check(self._L.wpic_ic_weibel2d(self._h, int(seed), int(n0), float(v_thi), float(v_the), float(t_ani),
                                       float(b0)))

    def ic_harris(self, seed, nbg, ncs, lcs, x0, y0, b0, e1, rtemp, vti, vte):
        """the reference reconnection app's Harris sheet (2d/proj/reconnection/app.f90:260-452)"""
        check(self._L.wpic_ic_harris2d(self._h, int(seed), int(nbg), int(ncs), float(lcs), float(x0), float(y0),
                                       float(b0), float(e1), float(rtemp), float(vti), float(vte)))

    def set_cg_tol(self, err):
        """CG relative tolerance (the reference's err = 1e-6, 2d/common/field.f90:339)"""
        check(self._L.wpic_set_cg_tol2d(self._h, float(err)))

    def perturb_bz(self, amp, mode=1):
        check(self._L.wpic_perturb_bz2d(self._h, float(amp), int(mode)))

    def adj_forward(self, nsteps, every):
        """nsteps with a checkpoint every `every` steps; returns the final W_B"""
        wb = C.c_double()
        check(self._L.wpic_adj_forward2d(self._h, int(nsteps), int(every), C.byref(wb)))
        return wb.value

    def adj_backward(self, seed=1.0):
        check(self._L.wpic_adj_backward2d(self._h, float(seed)))

    def adj_param_grads(self, udrift, uth_e, bz_mode=1):
        out = np.zeros(3)
        check(self._L.wpic_adj_param_grads2d(self._h, float(udrift), float(uth_e), int(bz_mode), out))
        return out

    def adj_download(self, np_cap):
        lam_up = np.zeros((6, np_cap, self.ny, self.nsp), order="F")
        lam_uf = np.zeros((6, self.nx + 4, self.ny + 4), order="F")
        check(self._L.wpic_adj_download2d(self._h, lam_up, lam_uf, np_cap))
        return lam_up, lam_uf

    def set_timing(self, on=True):
        check(self._L.wpic_set_timing(self._h, 1 if on else 0))

    def timing(self, name):
        ms, cnt = C.c_double(), C.c_longlong()
        check(self._L.wpic_timing(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


def weibel_beams_2d(nx, ny, n_ppc, *, seed=20261202, omega_pe=0.1, mass_ratio=25.0, udrift=0.5,
                    uth_e=0.1, uth_i=0.01, delt=0.5, mode=MODE_FAST, capacity=None):
    """The 2D Weibel beam workload: two species (ions first), periodic box,
    dx = c = 1, seeded co-located random in-cell positions, electron beams at
    u_x = +/-udrift with thermal spread uth_e, cold ions at uth_i."""
    q, r = weibel_charges(n_ppc, omega_pe, mass_ratio)
    cap = capacity if capacity is not None else n_ppc * nx * ny
    sim = Pic2D(nx, ny, q, r, delt=delt, mode=mode, capacity=cap)
    sim.ic_beams(seed, n_ppc, udrift, uth_e, uth_i, 0.0)
    return sim
