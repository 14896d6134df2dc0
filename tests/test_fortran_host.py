"""The Fortran host layer (fortnet_amd/fortran: the ISO_C_BINDING module, the host types and the
fnet_fortran_check driver, built with amdflang) drives the C-ABI the way the reference's Fortran
would: the reference's own fixtures must come out of it.

* CPU: without a usable device the driver stops with the library's message and a nonzero
  status, the reference's error() convention (message.F90:73-84).
* GPU: the analytic-derivative fixtures (SiC cells, G1-G5, species-resolved and unresolved,
  one and several targets) give the stored raw predictions and forces at the reference's
  tolerance (testwithworkdir.py:24-25); the steepest-descent restart fixtures give the written
  parameters after one gradient + update step.
"""
import os
import subprocess

import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import build, dataset
from conftest import case_names, load_case

DRIVER = build.FCHECK
ATOL, RTOL = 1e-10, 1e-9  # testwithworkdir.py:24-25

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER),
                                  reason="Fortran driver not built (no amdflang)")


def write_case(path, ns, geos, precision, train=None):
    """Case file of fnet_fortran_check (format in fnet_fortran_check.f90)."""
    g = ns.g
    L = [str(int(precision)), str(len(g))]
    for f in range(len(g)):
        L.append("%d %.17g %.17g %.17g %.17g %.17g %.17g %d %d %d" % (
            g.type[f], g.rcut[f], g.eta[f], g.rs[f], g.kappa[f], g.lam[f], g.xi[f], g.atomid[f],
            g.z1[f], g.z2[f]))
    if ns.zmean is not None and ns.zstd is not None:
        L += ["1", " ".join("%.17g" % x for x in ns.zmean), " ".join("%.17g" % x for x in ns.zstd)]
    else:
        L.append("0")
    ext_in = [int(k) for k in (ns.ext_in if ns.ext_in is not None else [])]
    L.append(str(len(ext_in)))
    if ext_in:
        L.append(" ".join(str(k) for k in ext_in))
    L += [str(len(ns.species_z)), " ".join(str(int(z)) for z in ns.species_z),
          " ".join(str(int(a)) for a in ns.activation), str(len(ns.dims)),
          " ".join(str(int(d)) for d in ns.dims), str(ns.params.shape[1])]
    L += ["%.17g" % x for x in ns.params.reshape(-1)]
    L.append(str(len(geos)))
    next_ = geos[0].ext.shape[1] if geos else 0
    L.append(str(next_))
    for geo in geos:
        L.append("%d %d" % (geo.n, 1 if geo.periodic else 0))
        L.append(" ".join("%.17g" % x for x in geo.latvecs.reshape(-1)))
        for a in range(geo.n):
            L.append(" ".join(["%.17g %.17g %.17g %d %d" % (tuple(geo.coords[a]) + (
                geo.atnum[a], geo.species[a] + 1))] + ["%.17g" % x for x in geo.ext[a]]))
    if train is None:
        L.append("0")
    else:
        nglob, gtargets, atargets, aw, dw, lr, maxdisp, tol = train
        L.append("1")
        L.append(str(int(nglob)))
        if nglob > 0:  # (nGlobal, nStruct) in Fortran order = [nStruct][nGlobal] rows
            L += ["%.17g" % x for x in np.asarray(gtargets, np.float64).reshape(-1)]
        if nglob < ns.n_out:  # (nOut - nGlobal, nAtom) in Fortran order = [nAtom][...] rows
            L += ["%.17g" % x for x in np.asarray(atargets, np.float64).reshape(-1)]
        L += ["%.17g" % x for x in aw]
        L += ["%.17g" % x for x in dw]
        L.append("%.17g %.17g %.17g" % (lr, maxdisp, tol))
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def read_result(path):
    """{name: array (n1, n2) in Fortran order} of a result file."""
    out = {}
    lines = open(path).read().split("\n")
    k = 0
    while k < len(lines) and lines[k].strip():
        name, n1, n2 = lines[k].split()
        n1, n2 = int(n1), int(n2)
        vals = np.array([float(x) for x in lines[k + 1:k + 1 + n1 * n2]])
        out[name] = vals.reshape(n2, n1).T
        k += 1 + n1 * n2
    return out


def run_driver(case, result, device=0):
    return subprocess.run([DRIVER, str(case), str(result), str(device)], capture_output=True,
                          text=True, timeout=300)


def _tiny():
    g = fa.GFunctions([2, 5], [8.0, 8.0], eta=[0.1, 0.01], xi=[0.0, 1.0], lam=[0.0, 1.0])
    dims = [2, 3, 1]
    ns = fa.Netstat(g, [14], dims, 0, np.zeros((1, fa.nparams(dims))))
    geo = fa.Geometry(np.array([[0.0, 0.0, 0.0], [4.0, 0.0, 0.0], [0.0, 4.0, 0.0]]),
                      [14, 14, 14], [0, 0, 0], np.eye(3) * 10.0)
    return ns, [geo]


@needs_driver
def test_driver_stops_with_message_without_device(tmp_path):
    ns, geos = _tiny()
    write_case(tmp_path / "case.txt", ns, geos, fa.FNET_FP64)
    r = run_driver(tmp_path / "case.txt", tmp_path / "out.txt", device=999)
    assert r.returncode != 0
    assert "ERROR! fnet_ctx_create failed" in r.stderr
    assert not (tmp_path / "out.txt").exists()


FORCE_CASES = [n for n in case_names() if n.startswith("analysis__derivatives__analytical")
               and "atomID" not in n]
# every steepest-descent restart: global, atomic and both kinds of targets
SD_CASES = [n for n in case_names("restart") if str(load_case(n)[0]["optimizer"]) == "sd"]
# predictions with external atomic features beside the ACSF (mixed) or an empty external list
EXT_CASES = ["input__mixedfeatures__prediction", "input__mixedfeatures__validation",
             "input__extfeatures__prediction__empty", "input__extfeatures__validation__empty"]


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("name", FORCE_CASES)
def test_fortran_host_reproduces_force_fixture(name, tmp_path):
    c, ds = load_case(name)
    print("case", name, flush=True)
    ns = fa.Netstat.from_arrays(c, "in_")
    geos = dataset.geometries(ds)
    write_case(tmp_path / "case.txt", ns, geos, fa.FNET_FP64)
    r = run_driver(tmp_path / "case.txt", tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    res = read_result(tmp_path / "out.txt")
    for got, want in ((res["A"].T, c["out_raw"]), (res["F"].T, c["out_forces"])):
        # element-wise, as testwithworkdir.py:700-714 compares
        assert np.allclose(got, want, atol=ATOL, rtol=RTOL), np.abs(got - want).max()
    assert res["G"].shape == (len(ns.g), sum(g.n for g in geos))


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("name", SD_CASES)
def test_fortran_host_sd_step_matches_fixture(name, tmp_path):
    """One gradient + steepest-descent step from the restart netstat, driven from Fortran, gives
    the netstat the reference wrote (sysTrain/bprop and steepdesc.F90:160-204)."""
    c, ds = load_case(name)
    ns = fa.Netstat.from_arrays(c, "in_")
    print("case", name, flush=True)
    geos = dataset.geometries(ds)
    maxdisp = float(c["maxdisp"]) if float(c["maxdisp"]) > 0 else 1e4
    ng = ns.n_global
    train = (ng, ds["globaltargets"] if ng > 0 else None,
             ds["atomictargets"] if ng < ns.n_out else None, ds["atomicweights"],
             ds["weights"].astype(np.float64), float(c["learningrate"]), maxdisp, 1e-20)
    write_case(tmp_path / "case.txt", ns, geos, fa.FNET_FP64, train)
    r = run_driver(tmp_path / "case.txt", tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    res = read_result(tmp_path / "out.txt")
    new = res["PARAMS"].T  # [nSpecies, nParams]
    assert np.allclose(new, c["after_params"], atol=ATOL, rtol=RTOL), \
        np.abs(new - c["after_params"]).max()


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("name", EXT_CASES)
def test_fortran_host_external_features(name, tmp_path):
    """External atomic features through the Fortran types (TFnetGeometry%ext, TFnetModel%extIn):
    the raw predictions of the reference's external-feature runs, element-wise."""
    c, ds = load_case(name)
    print("case", name, flush=True)
    ns = fa.Netstat.from_arrays(c, "in_")
    geos = dataset.geometries(ds)
    write_case(tmp_path / "case.txt", ns, geos, fa.FNET_FP64)
    r = run_driver(tmp_path / "case.txt", tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    res = read_result(tmp_path / "out.txt")
    got = res["A"].T
    assert np.allclose(got, c["out_raw"], atol=ATOL, rtol=RTOL), np.abs(got - c["out_raw"]).max()
    if "out_glob" in c and c["out_glob"].size:
        assert np.allclose(res["E"].T[:, :ns.n_global], c["out_glob"], atol=ATOL, rtol=RTOL)
