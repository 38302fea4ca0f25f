"""MI355X-native WumingPIC step loop: HIP kernels for gfx950 behind the C-ABI
of include/wpic.h (2D) and include/wpic3d.h (3D), with host-side mirrors of
the reference's module entry points (pic2d.Pic2D, pic3d.Pic3D) and an
ISO_C_BINDING Fortran shim (fortran/)."""
from ._lib import BC_PERIODIC, BC_SHOCK, BC_WALLX, MODE_FAST, MODE_MIXED, MODE_PARITY, WpicError  # noqa: F401

__all__ = ["BC_PERIODIC", "BC_WALLX", "MODE_FAST", "MODE_MIXED", "MODE_PARITY", "WpicError", "Pic2D", "Pic3D",
           "weibel_beams_2d", "weibel_beams_3d"]


def __getattr__(name):
    if name in ("Pic2D", "weibel_beams_2d", "weibel_charges"):
        from . import pic2d
        return getattr(pic2d, name)
    if name in ("Pic3D", "weibel_beams_3d"):
        from . import pic3d
        return getattr(pic3d, name)
    raise AttributeError(name)
