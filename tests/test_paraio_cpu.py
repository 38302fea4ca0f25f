"""The raw + JSON file pairs of include/wpic_io.h (host C++, no GPU) against
the reference's own jsonio / JSON-Fortran (golden file pairs in
tests/golden/ref2d_paraio.npz written by tests/golden/gen_ref_paraio.py, and
fresh cases through oracle/_ref when it is built), byte for byte; read back
the way python/json2hdf5.py reads them; paraio__input round trips and
errors."""
import json
import os
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import gen_ref_paraio as G  # noqa: E402

GOLD = os.path.join(HERE, "golden", "ref{}_paraio.npz")


def _io(P, d):
    from wumingpic_amd.paraio import ParaIO
    z = dict(nzgs=int(P["nzgs"]), nzge=int(P["nzge"])) if "nzgs" in P else {}
    return ParaIO(ndim=int(P["ndim"]), np_cap=int(P["np"]), nsp=int(P["nsp"]), nxgs=int(P["nxgs"]),
                  nxge=int(P["nxge"]), nygs=int(P["nygs"]), nyge=int(P["nyge"]), delx=float(P["delx"]),
                  delt=float(P["delt"]), c=float(P["c"]), q=np.asarray(P["q"]), r=np.asarray(P["r"]), dir=d + "/",
                  **z)


def _write_all(P, up, uf, np2, mom, d, routines=G.ROUTINES):
    io = _io(P, d)
    m = np.array(mom, order="F")
    it = int(P["it"])
    for rt in routines:
        if rt == "output":
            io.output(up, uf, np2, int(P["nxs"]), int(P["nxe"]), it, "restart")
        elif rt == "param":
            io.param(int(P["n0"]), *(float(P[k]) for k in ("wpe", "wpi", "wge", "wgi", "vti", "vte")), "param")
        elif rt == "mom":
            io.mom(m, uf, it)
        elif rt == "ptcl":
            io.ptcl(up, uf, np2, it)
        else:
            io.orb(up, uf, np2, it)
    names = G.file_names(P)
    out = {}
    for rt in routines:
        with open(os.path.join(d, names[rt] + ".json"), "rb") as f:
            js = f.read()
        with open(os.path.join(d, names[rt] + ".raw"), "rb") as f:
            raw = f.read()
        out[rt] = (js, raw)
    return out, m


def _gold(tag="2d"):
    d = np.load(GOLD.format(tag))
    P = {k: d[k] for k in ("ndim", "np", "nsp", "nxgs", "nxge", "nygs", "nyge", "it", "nxs", "nxe", "n0", "delx",
                           "delt", "c", "wpe", "wpi", "wge", "wgi", "vti", "vte", "q", "r")}
    if "nzgs" in d:
        P["nzgs"], P["nzge"] = d["nzgs"], d["nzge"]
    return d, P


@pytest.mark.parametrize("tag", ["2d", "3d"])
@pytest.mark.parametrize("routine", G.ROUTINES)
def test_file_pairs_match_reference_golden(routine, tag):
    d, P = _gold(tag)
    with tempfile.TemporaryDirectory() as t:
        out, m = _write_all(P, d["up"], d["uf"], d["np2"], d["mom"], t, (routine,))
    js, raw = out[routine]
    assert js == d[f"{routine}_json"].tobytes()
    assert raw == d[f"{routine}_raw"].tobytes()
    if routine == "mom":
        # paraio__mom divides mom(2:7) by mom(1) in place
        assert np.array_equal(m.view(np.uint64), d["mom_divided"].view(np.uint64))


def _read_like_json2hdf5(jsonfile):
    """python/json2hdf5.py:16-83: endian flag, attributes by offset, datasets
    with the shape reversed (column-major)"""
    with open(jsonfile) as f:
        obj = json.load(f)
    assert obj["meta"]["endian"] == 1
    rawfile = os.path.join(os.path.dirname(jsonfile), obj["meta"]["rawfile"])
    raw = open(rawfile, "rb").read()
    out = {}
    for grp in ("attribute", "dataset"):
        for name, m in obj.get(grp, {}).items():
            dt = np.dtype("<" + m["datatype"])
            n = int(np.prod(m["shape"]))
            assert m["size"] == n * dt.itemsize
            a = np.frombuffer(raw, dtype=dt, count=n, offset=m["offset"]).reshape(m["shape"][::-1])
            out[name] = a
            if grp == "attribute":
                data = np.atleast_1d(np.array(m["data"]))
                assert np.array_equal(a.ravel().astype(data.dtype), data.ravel()), name
    return out


def test_files_read_back_as_json2hdf5_reads_them():
    d, P = _gold()
    up, uf, np2, mom = d["up"], d["uf"], d["np2"], d["mom"]
    nx, ny = int(P["nxge"] - P["nxgs"] + 1), int(P["nyge"] - P["nygs"] + 1)
    with tempfile.TemporaryDirectory() as t:
        out, m = _write_all(P, up, uf, np2, mom, t)
        r = _read_like_json2hdf5(os.path.join(t, "restart.json"))
        assert np.array_equal(r["up"].reshape(-1).view(np.uint64), up.ravel(order="F").view(np.uint64))
        assert np.array_equal(r["np2"].reshape(-1), np2.ravel(order="F"))
        assert np.array_equal(r["uf"].reshape(-1), uf.ravel(order="F"))
        assert int(r["it"][0]) == int(P["it"]) and int(r["nsp"][0]) == int(P["nsp"])
        mm = _read_like_json2hdf5(os.path.join(t, "%07d_mom.json" % int(P["it"])))
        # den[s, j, i] = mom(1, i, j, s) of the interior nodes
        inner = mom[:, 1:nx + 1, 1:ny + 1, :]
        assert np.array_equal(mm["den"], inner[0].transpose(2, 1, 0))
        with np.errstate(divide="ignore", invalid="ignore"):
            vel = inner[1:4] / inner[0:1]
        assert np.allclose(mm["vel"], vel.transpose(3, 2, 1, 0), equal_nan=True)
        pt = _read_like_json2hdf5(os.path.join(t, "%07d_ptcl.json" % int(P["it"])))
        ob = _read_like_json2hdf5(os.path.join(t, "%07d_orb.json" % int(P["it"])))
        for s in range(int(P["nsp"])):
            act = np.concatenate([up[:, :np2[j, s], j, s] for j in range(np2.shape[0])], axis=1)
            assert np.array_equal(pt["up%02d" % (s + 1)].view(np.uint64), act.T.view(np.uint64))
            ids = act[-1].view(np.int64)
            assert np.array_equal(ob["up%02d" % (s + 1)].view(np.uint64), act[:, ids > 0].T.view(np.uint64))


def test_output_then_input_round_trip_and_grid_check():
    from wumingpic_amd._lib import WpicError
    d, P = _gold()
    up, uf, np2 = d["up"], d["uf"], d["np2"]
    with tempfile.TemporaryDirectory() as t:
        io = _io(P, t)
        io.output(up, uf, np2, 3, 6, 77, "snap")
        up2 = np.zeros_like(up, order="F")
        uf2 = np.zeros_like(uf, order="F")
        n2 = np.zeros_like(np2, order="F")
        nxs, nxe, it = io.input(up2, uf2, n2, "snap")
        assert (nxs, nxe, it) == (3, 6, 77)
        assert np.array_equal(up2.view(np.uint64), up.view(np.uint64))
        assert np.array_equal(uf2, uf) and np.array_equal(n2, np2)
        # a file of another grid is refused with the reference's message
        P2 = dict(P)
        P2["nxge"] = P["nxge"] + 1
        io2 = _io(P2, t)
        with pytest.raises(WpicError, match="invalid input parameters"):
            io2.input(np.zeros_like(up, order="F"), np.zeros((6, int(P2["nxge"] - P2["nxgs"] + 5), uf.shape[2]),
                                                             order="F"), n2, "snap")


@pytest.mark.parametrize("tag", ["2d", "3d"])
def test_input_reads_the_reference_written_restart(tag):
    d, P = _gold(tag)
    with tempfile.TemporaryDirectory() as t:
        with open(os.path.join(t, "restart.json"), "wb") as f:
            f.write(d["output_json"].tobytes())
        with open(os.path.join(t, "restart.raw"), "wb") as f:
            f.write(d["output_raw"].tobytes())
        io = _io(P, t)
        up2 = np.zeros_like(d["up"], order="F")
        uf2 = np.zeros_like(d["uf"], order="F")
        n2 = np.zeros_like(d["np2"], order="F")
        nxs, nxe, it = io.input(up2, uf2, n2, "restart")
    assert (nxs, nxe, it) == (int(P["nxs"]), int(P["nxe"]), int(P["it"]))
    assert np.array_equal(uf2, d["uf"]) and np.array_equal(n2, d["np2"])
    if tag == "2d":
        # the padded particle array, inactive slots included
        assert np.array_equal(up2.view(np.uint64), d["up"].view(np.uint64))
    else:
        # 3D files hold the active particles only (3d/common/paraio.f90:206-231)
        ny, nz, nsp = d["np2"].shape
        for s in range(nsp):
            for k in range(nz):
                for j in range(ny):
                    n = d["np2"][j, k, s]
                    assert np.array_equal(up2[:, :n, j, k, s].view(np.uint64), d["up"][:, :n, j, k, s].view(np.uint64))


def test_nonfinite_attributes_written_as_gfortran_does():
    """JSON-Fortran writes a non-finite real with the compiler's list-directed
    text in quotes; the reference's gcc build (compiler-gcc.mk) gives
    "Infinity", "-Infinity" and "NaN".  (The amdflang oracle build gives "Inf";
    this text is compiler-dependent and so parity unpinned by a build here.)"""
    d, P = _gold()
    with tempfile.TemporaryDirectory() as t:
        io = _io(P, t)
        io.param(16, float("inf"), 0.02, float("-inf"), 0.0, float("nan"), 0.1, "param")
        obj = json.load(open(os.path.join(t, "param.json")))
    a = obj["attribute"]
    assert a["wpe"]["data"] == "Infinity" and a["wge"]["data"] == "-Infinity" and a["vti"]["data"] == "NaN"


@pytest.mark.skipif(not os.path.exists(G.DRIVER), reason="oracle/_ref not built (needs the reference checkout)")
def test_real_formatting_against_the_reference_library():
    """finite reals across the whole double range (zero, -0, subnormals,
    extremes, integers, random mantissas) through paraio__param: the JSON text
    equals what the reference's JSON-Fortran prints"""
    rng = np.random.default_rng(99)

    def rnd():
        k = rng.integers(0, 5)
        if k == 0:
            return float(rng.choice([0.0, -0.0, 1.0, -1.0, 0.5, 0.1, 1 / 3, 5e-324, -5e-324, 2.2250738585072014e-308,
                                     1.7976931348623157e308, 1e22, 123456789012345678.0]))
        if k == 1:
            return float(rng.integers(-10 ** 9, 10 ** 9))
        if k == 2:
            return float(np.float64(rng.integers(1, 2 ** 53)) * 2.0 ** int(rng.integers(-1074, 971)))
        if k == 3:
            return float(rng.standard_normal() * 10.0 ** int(rng.integers(-300, 300)))
        return float(rng.standard_normal() * 10.0 ** int(rng.integers(-12, 12)))

    d, P0 = _gold()
    for case in range(60):
        P = dict(P0)
        for k in ("delx", "delt", "c", "wpe", "wpi", "wge", "wgi", "vti", "vte"):
            P[k] = rnd()
        P["q"] = np.array([rnd(), rnd()])
        P["r"] = np.array([rnd(), rnd()])
        P["n0"] = int(rng.integers(-2 ** 31, 2 ** 31))
        with np.errstate(all="ignore"):
            derived = [np.float64(P["c"]) * P["wgi"] / P["wpi"], np.float64(P["c"]) * P["wge"] / P["wpe"]]
        if not np.all(np.isfinite(derived)):
            continue  # vai, vae non-finite: compiler-dependent text (see the test above)
        ref, _ = G.run_reference(P, d["up"], d["uf"], d["np2"], d["mom"], ("param",))
        with tempfile.TemporaryDirectory() as t:
            mine, _ = _write_all(P, d["up"], d["uf"], d["np2"], d["mom"], t, ("param",))
        assert mine["param"] == ref["param"], case


@pytest.mark.skipif(not os.path.exists(G.DRIVER3D), reason="oracle/_ref not built (needs the reference checkout)")
@pytest.mark.parametrize("tag", ["2d", "3d"])
def test_fresh_state_all_routines_against_the_reference_library(tag):
    P, up, uf, np2, mom = (G.case_inputs if tag == "2d" else G.case_inputs3d)(seed=5)
    ref, divided = G.run_reference(P, up, uf, np2, mom)
    with tempfile.TemporaryDirectory() as t:
        mine, m = _write_all(P, up, uf, np2, mom, t)
    for rt in G.ROUTINES:
        assert mine[rt] == ref[rt], rt
    assert np.array_equal(m.view(np.uint64), divided.view(np.uint64))


def test_shock_param_appends_u0():
    """the shock app's save_param (2d/proj/shock/app.f90:567-612) re-prints paraio__param's JSON
    with the attribute u0 added last, its 8 bytes appended to the raw file: everything before it is
    paraio__param's pair byte for byte"""
    import json as _json
    import tempfile as _tf
    from wumingpic_amd.paraio import ParaIO
    with _tf.TemporaryDirectory() as t:
        io = ParaIO(ndim=6, np_cap=40, nsp=2, nxgs=2, nxge=9, nygs=2, nyge=5, delx=1.0, delt=1.0, c=1.0,
                    q=[0.01, -0.01], r=[25.0, 1.0], dir=t + "/")
        io.param(4, 0.1, 0.02, 0.03, 0.0012, 0.01, 0.05, "a")
        io.param_shock(4, 0.1, 0.02, 0.03, 0.0012, 0.01, 0.05, -0.3, "b")
        io.close()
        ra, rb = (open(os.path.join(t, n + ".raw"), "rb").read() for n in "ab")
        ja, jb = (_json.load(open(os.path.join(t, n + ".json"))) for n in "ab")
        ta = open(os.path.join(t, "a.json")).read()
        tb = open(os.path.join(t, "b.json")).read()
    assert rb[:len(ra)] == ra and len(rb) == len(ra) + 8
    assert np.frombuffer(rb[-8:], dtype="<f8")[0] == -0.3
    assert list(jb["attribute"])[-1] == "u0" and jb["attribute"]["u0"]["offset"] == len(ra)
    assert list(ja["attribute"]) == list(jb["attribute"])[:-1]
    # the JSON text differs only by the new entry and the raw file name
    assert tb.replace('"b.raw"', '"a.raw"').startswith(ta[:ta.index('"n0"')])
