"""MI355X-native WumingPIC step loop: HIP kernels for gfx950 behind the C-ABI
of include/wpic.h, with a host-side mirror of the reference's module entry
points (pic2d.Pic2D) and an ISO_C_BINDING Fortran shim (fortran/)."""
from ._lib import BC_PERIODIC, BC_WALLX, MODE_FAST, MODE_PARITY, WpicError  # noqa: F401

__all__ = ["BC_PERIODIC", "BC_WALLX", "MODE_FAST", "MODE_PARITY", "WpicError", "Pic2D", "weibel_beams_2d"]


def __getattr__(name):
    if name in ("Pic2D", "weibel_beams_2d", "weibel_charges"):
        from . import pic2d
        return getattr(pic2d, name)
    raise AttributeError(name)
