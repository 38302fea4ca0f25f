"""Build the gfx950 shared library wumingpic_amd/libwpic.so from csrc/*.hip
(device code, hipcc) and csrc/*.cpp (host-only file formats, g++).

    python -m wumingpic_amd.build          (or build() from Python)

Every call recompiles (hipcc cross-compiles for gfx950 without a GPU).  The
FP64 path is compiled without FMA contraction (-ffp-contract=off) so its
arithmetic follows the reference's evaluation order operation for operation;
packed FP32 instructions are disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwpic.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


CXX = os.environ.get("CXX", "g++")
# host-only sources (file formats): plain C++
CXXFLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src.endswith(".cpp"):
        cmd = [CXX, *CXXFLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


FORT = os.path.join(HERE, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/amdflang")
FSRC = ["wpic_capi.f90", "wpic_shim_state.f90", "shim_modules.f90"]
FDRIVER = "shim_step2d_driver.f90"
FADJ = "adj_driver.f90"
# the 3d/common modules repeat the 2D module names: their own module
# directory, archive and driver
FSRC3D = ["wpic_capi3d.f90", "wpic_shim_state3d.f90", "shim_modules3d.f90"]
FDRIVER3D = "shim_step3d_driver.f90"


def build_fortran():
    """The ISO_C_BINDING shim modules (the reference's module names) as
    fortran/libwpic_fortran.a, and the app-loop driver linked to libwpic.so."""
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for f in FSRC + [FDRIVER]:
        obj = os.path.join(mod, f + ".o")
        cmd = [FLANG, "-O2", "-fPIC", "-module-dir", mod, "-I", mod, "-c", os.path.join(FORT, f), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {f}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    archive = os.path.join(FORT, "libwpic_fortran.a")
    if os.path.exists(archive):
        os.remove(archive)
    subprocess.run(["ar", "rcs", archive, *objs[:-1]], check=True)
    exe = os.path.join(FORT, "shim_step2d_driver")
    cmd = [FLANG, "-o", exe, objs[-1], archive, "-L" + HERE, "-lwpic", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of the Fortran driver failed:\n{r.stdout}\n{r.stderr}")
    # the adjoint called from Fortran
    aobj = os.path.join(mod, FADJ + ".o")
    cmd = [FLANG, "-O2", "-fPIC", "-module-dir", mod, "-I", mod, "-c", os.path.join(FORT, FADJ), "-o", aobj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for {FADJ}:\n{r.stdout}\n{r.stderr}")
    aexe = os.path.join(FORT, "adj_driver")
    cmd = [FLANG, "-o", aexe, aobj, archive, "-L" + HERE, "-lwpic", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of the Fortran adjoint driver failed:\n{r.stdout}\n{r.stderr}")
    build_fortran3d()
    return exe


def build_fortran3d():
    """The 3D shim modules (3d/common's module names) as
    fortran/libwpic_fortran3d.a, and the 3D app-loop driver."""
    mod = os.path.join(OBJ, "fmod3d")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for f in FSRC3D + [FDRIVER3D]:
        obj = os.path.join(mod, f + ".o")
        cmd = [FLANG, "-O2", "-fPIC", "-module-dir", mod, "-I", mod, "-c", os.path.join(FORT, f), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {f}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    archive = os.path.join(FORT, "libwpic_fortran3d.a")
    if os.path.exists(archive):
        os.remove(archive)
    subprocess.run(["ar", "rcs", archive, *objs[:-1]], check=True)
    exe = os.path.join(FORT, "shim_step3d_driver")
    cmd = [FLANG, "-o", exe, objs[-1], archive, "-L" + HERE, "-lwpic", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of the 3D Fortran driver failed:\n{r.stdout}\n{r.stderr}")
    return exe


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_app()
    build_fortran()
    return LIB


APPS = ("wpic_app2d", "wpic_app3d")


def build_app():
    """the command-line problem drivers, linked to libwpic.so"""
    for name in APPS:
        exe = os.path.join(HERE, "app", name)
        cmd = [CXX, *CXXFLAGS, "-o", exe, exe + ".cpp", "-L" + HERE, "-lwpic", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build of {name} failed:\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    print(build(verbose=True))
