"""Write golden file pairs of the reference's 2D and 3D paraio routines at
one rank (output, param, mom, ptcl, orb; 2d/common/paraio.f90,
3d/common/paraio.f90), produced through the reference's OWN jsonio module and
JSON-Fortran library by oracle/_ref/ref_paraio{2d,3d} (`make -C oracle ref`).

    python tests/golden/gen_ref_paraio.py

Output: tests/golden/ref{2d,3d}_paraio.npz with the inputs and, per file, the
bytes of the .json and .raw the reference library writes.  Data only.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_paraio2d")
DRIVER3D = os.path.join(ROOT, "oracle", "_ref", "ref_paraio3d")
ROUTINES = ("output", "param", "mom", "ptcl", "orb")


def case_inputs(seed=20261206):
    rng = np.random.default_rng(seed)
    P = dict(ndim=6, np=20, nsp=2, nxgs=2, nxge=7, nygs=2, nyge=5, it=1234, nxs=2, nxe=7, n0=16,
             delx=1.0, delt=0.5, c=1.0, wpe=0.1, wpi=0.02, wge=0.05, wgi=0.002, vti=0.01, vte=0.1,
             q=np.array([0.0123456789, -0.0123456789]), r=np.array([25.0, 1.0]))
    nx, ny = P["nxge"] - P["nxgs"] + 1, P["nyge"] - P["nygs"] + 1
    up = np.asfortranarray(rng.standard_normal((P["ndim"], P["np"], ny, P["nsp"])))
    # ids: raw 64-bit integers, some not positive (the orbit file keeps the
    # positive ones, paraio.f90:918-927)
    ids = rng.integers(-1000, 1000, size=(P["np"], ny, P["nsp"])).astype(np.int64)
    up[5] = ids.view(np.float64)
    uf = np.asfortranarray(rng.standard_normal((6, nx + 4, ny + 4)))
    np2 = np.asfortranarray(rng.integers(0, P["np"] + 1, size=(ny, P["nsp"])).astype(np.int32))
    mom = np.asfortranarray(rng.random((7, nx + 2, ny + 2, P["nsp"])) + 0.5)
    mom[0, 1, 1, 0] = 0.0  # a node without particles: 0/0 and x/0 in the division
    return P, up, uf, np2, mom


def case_inputs3d(seed=20261207):
    rng = np.random.default_rng(seed)
    P = dict(ndim=7, np=12, nsp=2, nxgs=2, nxge=5, nygs=2, nyge=4, nzgs=2, nzge=4, it=42, nxs=2, nxe=5, n0=8,
             delx=1.0, delt=0.5, c=1.0, wpe=0.1, wpi=0.02, wge=0.05, wgi=0.002, vti=0.01, vte=0.1,
             q=np.array([0.0123456789, -0.0123456789]), r=np.array([25.0, 1.0]))
    nx, ny, nz = (P["nxge"] - P["nxgs"] + 1, P["nyge"] - P["nygs"] + 1, P["nzge"] - P["nzgs"] + 1)
    up = np.asfortranarray(rng.standard_normal((P["ndim"], P["np"], ny, nz, P["nsp"])))
    ids = rng.integers(-1000, 1000, size=(P["np"], ny, nz, P["nsp"])).astype(np.int64)
    up[6] = ids.view(np.float64)
    uf = np.asfortranarray(rng.standard_normal((6, nx + 4, ny + 4, nz + 4)))
    np2 = np.asfortranarray(rng.integers(0, P["np"] + 1, size=(ny, nz, P["nsp"])).astype(np.int32))
    mom = np.asfortranarray(rng.random((7, nx + 2, ny + 2, nz + 2, P["nsp"])) + 0.5)
    mom[0, 1, 1, 1, 1] = 0.0
    return P, up, uf, np2, mom


def file_names(P):
    it = P["it"]
    return {"output": "restart", "param": "param", "mom": "%07d_mom" % it, "ptcl": "%07d_ptcl" % it,
            "orb": "%07d_orb" % it}


def run_reference(P, up, uf, np2, mom, routines=ROUTINES):
    """run the reference-library driver (2D or 3D by the parameters); return
    {name: (json bytes, raw bytes)} and the moments as paraio__mom leaves them"""
    out = {}
    three = "nzgs" in P
    keys = ("ndim", "np", "nsp", "nxgs", "nxge", "nygs", "nyge") + (("nzgs", "nzge") if three else ()) + (
        "it", "nxs", "nxe", "n0")
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "params.txt"), "w") as f:
            f.write(" ".join(str(P[k]) for k in keys) + "\n")
            f.write(" ".join(repr(float(P[k])) for k in ("delx", "delt", "c", "wpe", "wpi", "wge", "wgi", "vti",
                                                          "vte")) + "\n")
            f.write(" ".join(repr(float(v)) for v in P["q"]) + "\n")
            f.write(" ".join(repr(float(v)) for v in P["r"]) + "\n")
        for nm, a in (("up", up), ("uf", uf), ("mom", mom), ("np2", np2)):
            a.ravel(order="F").tofile(os.path.join(d, nm + ".bin"))
        names = file_names(P)
        for rt in routines:
            subprocess.run([DRIVER3D if three else DRIVER, d, rt], check=True)
            with open(os.path.join(d, names[rt] + ".json"), "rb") as f:
                js = f.read()
            with open(os.path.join(d, names[rt] + ".raw"), "rb") as f:
                raw = f.read()
            out[rt] = (js, raw)
        divided = None
        if "mom" in routines:
            divided = np.fromfile(os.path.join(d, "mom_divided.bin")).reshape(mom.shape, order="F")
    return out, divided


if __name__ == "__main__":
    if not (os.path.exists(DRIVER) and os.path.exists(DRIVER3D)):
        sys.exit("oracle/_ref/ref_paraio{2d,3d} are missing: run `make -C oracle ref` first")
    for tag, maker in (("2d", case_inputs), ("3d", case_inputs3d)):
        P, up, uf, np2, mom = maker()
        files, divided = run_reference(P, up, uf, np2, mom)
        arrays = {f"{rt}_json": np.frombuffer(js, dtype=np.uint8) for rt, (js, _) in files.items()}
        arrays.update({f"{rt}_raw": np.frombuffer(raw, dtype=np.uint8) for rt, (_, raw) in files.items()})
        scal = {k: np.array(v) for k, v in P.items()}
        out = os.path.join(HERE, f"ref{tag}_paraio.npz")
        np.savez_compressed(out, up=up, uf=uf, np2=np2, mom=mom, mom_divided=divided, **arrays, **scal)
        print("wrote", out)
