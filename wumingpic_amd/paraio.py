"""The reference's raw + JSON file pairs at one rank (paraio,
2d/common/paraio.f90 and 3d/common/paraio.f90) through the host C-ABI of
include/wpic_io.h.  Arrays are the reference's Fortran layouts (order='F')."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


class ParaIO:
    """paraio__init's module state; dir is prefixed to file names verbatim."""

    def __init__(self, *, ndim, np_cap, nsp, nxgs, nxge, nygs, nyge, delx, delt, c, q, r, dir,
                 nzgs=None, nzge=None):
        self._L = _lib.lib()
        self._h = C.c_void_p()
        q = np.ascontiguousarray(q, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        self.dims = 2 if nzgs is None else 3
        if self.dims == 2:
            check(self._L.wpic_paraio_init2d(C.byref(self._h), ndim, np_cap, nsp, nxgs, nxge, nygs, nyge, nygs,
                                             nyge, 1, 0, delx, delt, c, q, r, dir.encode()))
        else:
            check(self._L.wpic_paraio_init3d(C.byref(self._h), ndim, np_cap, nsp, nxgs, nxge, nygs, nyge, nzgs,
                                             nzge, nygs, nyge, nzgs, nzge, 1, 0, delx, delt, c, q, r,
                                             dir.encode()))

    def close(self):
        if self._h:
            self._L.wpic_paraio_finalize(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def output(self, up, uf, np2, nxs, nxe, it, filename):
        """paraio__output: 2d/common/paraio.f90:93, 3d/common/paraio.f90:102"""
        check(self._L.wpic_paraio_output(self._h, np.asfortranarray(up), np.asfortranarray(uf),
                                         np.asfortranarray(np2, dtype=np.int32), nxs, nxe, it, filename.encode()))

    def input(self, up, uf, np2, filename):
        """paraio__input into the given arrays; returns (nxs, nxe, it)"""
        nxs, nxe, it = C.c_int(), C.c_int(), C.c_int()
        check(self._L.wpic_paraio_input(self._h, up, uf, np2, C.byref(nxs), C.byref(nxe), C.byref(it),
                                        filename.encode()))
        return nxs.value, nxe.value, it.value

    def param(self, n0, wpe, wpi, wge, wgi, vti, vte, filename):
        """paraio__param: 2d/common/paraio.f90:431"""
        check(self._L.wpic_paraio_param(self._h, n0, wpe, wpi, wge, wgi, vti, vte, filename.encode()))

    def param_shock(self, n0, wpe, wpi, wge, wgi, vti, vte, u0, filename):
        """the shock app's save_param: paraio__param's pair with the attribute u0 appended"""
        check(self._L.wpic_paraio_param_shock(self._h, n0, wpe, wpi, wge, wgi, vti, vte, u0, filename.encode()))

    def mom(self, mom, uf, it):
        """paraio__mom: divides mom(2:7) by mom(1) in place, as the reference"""
        check(self._L.wpic_paraio_mom(self._h, mom, np.asfortranarray(uf), it))

    def ptcl(self, up, uf, np2, it):
        check(self._L.wpic_paraio_ptcl(self._h, np.asfortranarray(up), np.asfortranarray(uf),
                                       np.asfortranarray(np2, dtype=np.int32), it))

    def orb(self, up, uf, np2, it):
        check(self._L.wpic_paraio_orb(self._h, np.asfortranarray(up), np.asfortranarray(uf),
                                      np.asfortranarray(np2, dtype=np.int32), it))
