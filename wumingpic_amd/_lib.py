"""ctypes binding of libwpic.so (the C-ABI declared in include/wpic.h and
include/wpic3d.h).

The library is loaded from this package directory only; there is no CPU
fallback: importing the bindings without the built library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwpic.so")

MAXSP = 4
BC_PERIODIC, BC_WALLX, BC_SHOCK = 0, 1, 2
MODE_FAST, MODE_PARITY, MODE_MIXED = 0, 1, 2
PUSHER_BORIS, PUSHER_VAY = 0, 1
E_OK, E_INVAL, E_HIP, E_CG, E_MEM, E_DOMAIN = 0, -1, -2, -3, -4, -5

# every symbol include/wpic.h declares
EXPORTS = (
    "wpic_create2d", "wpic_destroy", "wpic_last_error", "wpic_set_mode", "wpic_sync",
    "wpic_upload_up2d", "wpic_download_up2d", "wpic_upload_gp2d", "wpic_download_gp2d",
    "wpic_upload_fields2d", "wpic_download_fields2d", "wpic_download_current2d",
    "wpic_download_dfield2d", "wpic_upload_dfield2d",
    "wpic_particle_solv2d", "wpic_particle_solv_vay2d", "wpic_set_pusher2d", "wpic_field_fdtd_i2d",
    "wpic_bc_particle_x2d", "wpic_bc_particle_y2d", "wpic_bc_injection2d", "wpic_set_injection2d",
    "wpic_inject2d", "wpic_ic_shock2d", "wpic_relocate2d", "wpic_set_nxe2d", "wpic_get_nxe2d",
    "wpic_sort_bucket2d", "wpic_cg_iters2d", "wpic_step2d",
    "wpic_energy2d", "wpic_gauss2d", "wpic_gauss2d_noalloc", "wpic_counts2d", "wpic_ic_beams2d",
    "wpic_set_timing", "wpic_timing",
    "wpic_mom_accl2d", "wpic_mom_nvt2d", "wpic_bc_mom2d", "wpic_moments2d",
    "wpic_download_mom2d", "wpic_upload_mom2d", "wpic_ic_weibel2d", "wpic_ic_harris2d",
)

# every symbol include/wpic3d.h declares
EXPORTS3D = (
    "wpic_create3d", "wpic_destroy3d", "wpic_set_mode3d", "wpic_sync3d",
    "wpic_upload_up3d", "wpic_download_up3d", "wpic_upload_gp3d", "wpic_download_gp3d",
    "wpic_upload_fields3d", "wpic_download_fields3d", "wpic_download_current3d",
    "wpic_download_dfield3d", "wpic_upload_dfield3d",
    "wpic_particle_solv3d", "wpic_particle_solv_vay3d", "wpic_set_pusher3d", "wpic_field_fdtd_i3d", "wpic_bc_particle_x3d", "wpic_bc_particle_yz3d",
    "wpic_sort_bucket3d", "wpic_cg_iters3d", "wpic_set_cg_tol3d", "wpic_step3d",
    "wpic_energy3d", "wpic_energy_species3d", "wpic_gauss3d", "wpic_counts3d", "wpic_ic_beams3d",
    "wpic_ic_weibel3d",
    "wpic_set_timing3d", "wpic_timing3d",
    "wpic_mom_accl3d", "wpic_mom_nvt3d", "wpic_bc_mom3d", "wpic_moments3d",
    "wpic_download_mom3d", "wpic_upload_mom3d",
)


# every symbol include/wpic_io.h declares
EXPORTS_IO = (
    "wpic_paraio_init2d", "wpic_paraio_init3d", "wpic_paraio_finalize", "wpic_paraio_output",
    "wpic_paraio_input", "wpic_paraio_param", "wpic_paraio_param_shock", "wpic_paraio_mom", "wpic_paraio_ptcl",
    "wpic_paraio_orb",
)


# every symbol include/wpic_app.h declares
EXPORTS_APP = ("wpic_app2d_main", "wpic_app3d_main")


# every symbol include/wpic_adj.h declares
EXPORTS_ADJ = (
    "wpic_set_cg_tol2d", "wpic_perturb_bz2d", "wpic_adj_forward2d", "wpic_adj_backward2d",
    "wpic_adj_param_grads2d", "wpic_adj_download2d",
)


class Params2D(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("nsp", C.c_int),
        ("nxgs", C.c_int), ("nxge", C.c_int), ("nygs", C.c_int), ("nyge", C.c_int),
        ("delx", C.c_double), ("delt", C.c_double), ("c", C.c_double), ("gfac", C.c_double),
        ("q", C.c_double * MAXSP), ("r", C.c_double * MAXSP),
        ("bc", C.c_int), ("mode", C.c_int), ("capacity", C.c_longlong),
    ]


class Params3D(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("nsp", C.c_int),
        ("nxgs", C.c_int), ("nxge", C.c_int), ("nygs", C.c_int), ("nyge", C.c_int),
        ("nzgs", C.c_int), ("nzge", C.c_int),
        ("delx", C.c_double), ("delt", C.c_double), ("c", C.c_double), ("gfac", C.c_double),
        ("q", C.c_double * MAXSP), ("r", C.c_double * MAXSP),
        ("mode", C.c_int), ("capacity", C.c_longlong),
    ]


class WpicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load libwpic.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m wumingpic_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
    i = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
    P = C.POINTER(Params2D)
    sig = {
        "wpic_create2d": [C.POINTER(vp), P],
        "wpic_destroy": [vp],
        "wpic_set_mode": [vp, C.c_int],
        "wpic_sync": [vp],
        "wpic_upload_up2d": [vp, d, i, i, C.c_int],
        "wpic_download_up2d": [vp, d, i, i, C.c_int],
        "wpic_upload_gp2d": [vp, d, i, C.c_int],
        "wpic_download_gp2d": [vp, d, i, C.c_int],
        "wpic_upload_fields2d": [vp, d],
        "wpic_download_fields2d": [vp, d],
        "wpic_download_current2d": [vp, d],
        "wpic_download_dfield2d": [vp, d],
        "wpic_upload_dfield2d": [vp, d],
        "wpic_particle_solv2d": [vp],
        "wpic_particle_solv_vay2d": [vp],
        "wpic_set_pusher2d": [vp, C.c_int],
        "wpic_field_fdtd_i2d": [vp],
        "wpic_bc_particle_x2d": [vp],
        "wpic_bc_injection2d": [vp, C.c_double],
        "wpic_set_injection2d": [vp, C.c_double],
        "wpic_ic_shock2d": [vp, C.c_ulonglong, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_double],
        "wpic_inject2d": [vp, C.c_ulonglong, C.c_int, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                          C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double],
        "wpic_relocate2d": [vp, C.c_ulonglong, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_double, C.c_double],
        "wpic_set_nxe2d": [vp, C.c_int],
        "wpic_get_nxe2d": [vp, C.POINTER(C.c_int)],
        "wpic_bc_particle_y2d": [vp],
        "wpic_sort_bucket2d": [vp],
        "wpic_cg_iters2d": [vp, np.ctypeslib.ndpointer(np.int32)],
        "wpic_step2d": [vp, C.c_int],
        "wpic_energy2d": [vp, np.ctypeslib.ndpointer(np.float64)],
        "wpic_gauss2d": [vp, d],
        "wpic_gauss2d_noalloc": [vp, d],
        "wpic_counts2d": [vp, np.ctypeslib.ndpointer(np.int64)],
        "wpic_ic_beams2d": [vp, C.c_ulonglong, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double],
        "wpic_set_timing": [vp, C.c_int],
        "wpic_timing": [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)],
        "wpic_ic_weibel2d": [vp, C.c_ulonglong, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double],
        "wpic_ic_harris2d": [vp, C.c_ulonglong, C.c_int, C.c_int] + [C.c_double] * 8,
        "wpic_mom_accl2d": [vp],
        "wpic_mom_nvt2d": [vp],
        "wpic_bc_mom2d": [vp],
        "wpic_moments2d": [vp],
        "wpic_download_mom2d": [vp, d],
        "wpic_upload_mom2d": [vp, d],
        "wpic_set_cg_tol2d": [vp, C.c_double],
        "wpic_perturb_bz2d": [vp, C.c_double, C.c_int],
        "wpic_adj_forward2d": [vp, C.c_int, C.c_int, C.POINTER(C.c_double)],
        "wpic_adj_backward2d": [vp, C.c_double],
        "wpic_adj_param_grads2d": [vp, C.c_double, C.c_double, C.c_int, np.ctypeslib.ndpointer(np.float64)],
        "wpic_adj_download2d": [vp, d, d, C.c_int],
        "wpic_create3d": [C.POINTER(vp), C.POINTER(Params3D)],
        "wpic_destroy3d": [vp],
        "wpic_set_mode3d": [vp, C.c_int],
        "wpic_sync3d": [vp],
        "wpic_upload_up3d": [vp, d, i, i, C.c_int],
        "wpic_download_up3d": [vp, d, i, i, C.c_int],
        "wpic_upload_gp3d": [vp, d, i, C.c_int],
        "wpic_download_gp3d": [vp, d, i, C.c_int],
        "wpic_upload_fields3d": [vp, d],
        "wpic_download_fields3d": [vp, d],
        "wpic_download_current3d": [vp, d],
        "wpic_download_dfield3d": [vp, d],
        "wpic_upload_dfield3d": [vp, d],
        "wpic_particle_solv3d": [vp],
        "wpic_particle_solv_vay3d": [vp],
        "wpic_set_pusher3d": [vp, C.c_int],
        "wpic_field_fdtd_i3d": [vp],
        "wpic_bc_particle_x3d": [vp],
        "wpic_bc_particle_yz3d": [vp],
        "wpic_sort_bucket3d": [vp],
        "wpic_cg_iters3d": [vp, np.ctypeslib.ndpointer(np.int32)],
        "wpic_set_cg_tol3d": [vp, C.c_double],
        "wpic_step3d": [vp, C.c_int],
        "wpic_energy3d": [vp, np.ctypeslib.ndpointer(np.float64)],
        "wpic_gauss3d": [vp, d],
        "wpic_counts3d": [vp, np.ctypeslib.ndpointer(np.int64)],
        "wpic_ic_beams3d": [vp, C.c_ulonglong, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double],
        "wpic_ic_weibel3d": [vp, C.c_ulonglong, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double],
        "wpic_energy_species3d": [vp, np.ctypeslib.ndpointer(np.float64)],
        "wpic_set_timing3d": [vp, C.c_int],
        "wpic_timing3d": [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)],
        "wpic_mom_accl3d": [vp],
        "wpic_mom_nvt3d": [vp],
        "wpic_bc_mom3d": [vp],
        "wpic_moments3d": [vp],
        "wpic_download_mom3d": [vp, d],
        "wpic_upload_mom3d": [vp, d],
    }
    ci = C.c_int
    cd = C.c_double
    dv = np.ctypeslib.ndpointer(dtype=np.float64)
    ip = C.POINTER(C.c_int)
    sig.update({
        "wpic_paraio_init2d": [C.POINTER(vp), ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, cd, cd, cd, dv, dv,
                               C.c_char_p],
        "wpic_paraio_init3d": [C.POINTER(vp), ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, cd, cd,
                               cd, dv, dv, C.c_char_p],
        "wpic_paraio_finalize": [vp],
        "wpic_paraio_output": [vp, d, d, i, ci, ci, ci, C.c_char_p],
        "wpic_paraio_input": [vp, d, d, i, ip, ip, ip, C.c_char_p],
        "wpic_paraio_param": [vp, ci, cd, cd, cd, cd, cd, cd, C.c_char_p],
        "wpic_paraio_param_shock": [vp, ci, cd, cd, cd, cd, cd, cd, cd, C.c_char_p],
        "wpic_paraio_mom": [vp, d, d, ci],
        "wpic_paraio_ptcl": [vp, d, d, i, ci],
        "wpic_paraio_orb": [vp, d, d, i, ci],
    })
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.wpic_last_error.restype = C.c_char_p
    L.wpic_last_error.argtypes = []
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise WpicError(rc, lib().wpic_last_error().decode())
    return rc
