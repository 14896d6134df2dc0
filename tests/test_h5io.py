"""HDF5 layer (include/fnet_h5.h over the HDF5 C library, fortnet_amd/h5io.py) against files the
reference wrote (tests/golden/hdf5, copied by tests/golden/copy_hdf5_fixtures.py).

CPU:
* every fnetdata dataset of the reference's tests reads (readHdfDataset, fnetdata.F90:994-1216)
  to exactly the arrays h5py gave for it (the converted .npz);
* every input and after-training netstat and every fnetout reads to the converted arrays;
* a netstat written from the after-training arrays of the steepest-descent restart cases has the
  tree of the reference's own _fortnet.hdf5 -- every group, attribute, dataset type, shape and
  value, dumped by h5py in a separate interpreter (tests/h5tree.py) -- and reads back unchanged;
* fnetout files written from the fixture outputs have the tree of the reference's _fnetout.hdf5
  (one fixture holds an empty atomictargets dataset that fnetout.F90:117-119 does not write when
  there are no atomic targets; it is the only difference allowed);
* the reference's error messages for invalid modes and netstat contents.
GPU:
* predict(): dataset + netstat files in, fnetout out, at the reference's tolerance against its
  _fnetout.hdf5 (testwithworkdir.py:24-25), for the analytic-force cases.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import h5io
from conftest import load_case

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H5 = os.path.join(GOLD, "hdf5")
ATOL, RTOL = 1e-10, 1e-9  # testwithworkdir.py:24-25
H5PY = "/opt/conda/bin/python3.9"


def _lib_ok():
    try:
        h5io._load()
        return True
    except fa.FnetError:
        return False


needs_lib = pytest.mark.skipif(not _lib_ok(), reason="libfnet_h5.so not built or HDF5 missing")
pytestmark = needs_lib

DATASETS = sorted(os.listdir(os.path.join(H5, "datasets")))
CASES = sorted(d for d in os.listdir(H5) if d != "datasets")
FNETOUT_CASES = [c for c in CASES if os.path.exists(os.path.join(H5, c, "_fnetout.hdf5"))]
SD_CASES = [c for c in CASES if c.startswith("restart")]


def _h5py_ok():
    if not os.path.exists(H5PY):
        return False
    r = subprocess.run([H5PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


needs_h5py = pytest.mark.skipif(not _h5py_ok(), reason="no interpreter with h5py")


def tree(path):
    r = subprocess.run([H5PY, os.path.join(os.path.dirname(__file__), "h5tree.py"), str(path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _equal_arrays(got, want):
    got, want = np.asarray(got), np.asarray(want)
    return got.shape == want.shape and np.array_equal(got, want)


@pytest.mark.parametrize("name", DATASETS)
def test_dataset_reads_as_h5py_does(name):
    d = h5io.read_dataset(os.path.join(H5, "datasets", name))
    ref = dict(np.load(os.path.join(GOLD, "datasets", name.replace(".hdf5", ".npz"))))
    assert set(ref) <= set(d)
    for k, v in ref.items():
        assert _equal_arrays(d[k], v), k


@pytest.mark.parametrize("case", CASES)
def test_netstat_and_fnetout_read_as_h5py_does(case):
    c, _ = load_case(case)
    for fn, pre in (("fortnet.hdf5", "in_"), ("_fortnet.hdf5", "after_")):
        p = os.path.join(H5, case, fn)
        if not os.path.exists(p):
            continue
        d = h5io.read_netstat(p)
        keys = [k for k in c if k.startswith(pre)]
        assert keys
        for k in keys:
            assert _equal_arrays(d[k[len(pre):]], c[k]), (fn, k)
        ns = h5io.load_netstat(p)
        assert ns.params.shape == c[pre + "params"].shape
    p = os.path.join(H5, case, "_fnetout.hdf5")
    if os.path.exists(p):
        o = h5io.read_fnetout(p)
        assert _equal_arrays(o["raw"], c["out_raw"])
        if "out_glob" in c and c["out_glob"].size:
            assert _equal_arrays(o["glob"], c["out_glob"])
        if int(c["out_tforces"]):
            assert _equal_arrays(o["forces"], c["out_forces"])


@pytest.mark.parametrize("case", SD_CASES)
def test_netstat_round_trip(case, tmp_path):
    c, _ = load_case(case)
    ns = fa.Netstat.from_arrays(c, "after_")
    p = tmp_path / "fortnet.hdf5"
    h5io.write_netstat(p, ns, n_atomic=int(c["after_natomictargets"]))
    d = h5io.read_netstat(p)
    for k in [k for k in c if k.startswith("after_")]:
        assert _equal_arrays(d[k[len("after_"):]], c[k]), k


@needs_h5py
@pytest.mark.parametrize("case", SD_CASES)
def test_written_netstat_has_the_reference_tree(case, tmp_path):
    """write_netstat of the arrays the reference stored after its steepest-descent step gives the
    reference's _fortnet.hdf5, group for group and value for value (netstat.F90:37-453,
    bpnn.F90:1062-1137, acsf.F90:1759-1904)."""
    c, _ = load_case(case)
    ns = fa.Netstat.from_arrays(c, "after_")
    p = tmp_path / "fortnet.hdf5"
    h5io.write_netstat(p, ns, n_atomic=int(c["after_natomictargets"]))
    assert tree(p) == tree(os.path.join(H5, case, "_fortnet.hdf5"))


def _write_case_fnetout(case, path, raw=None, forces=None):
    c, ds = load_case(case)
    ns = fa.Netstat.from_arrays(c, "in_")
    mode = h5io.read_fnetout(os.path.join(H5, case, "_fnetout.hdf5"))["mode"]
    ds = h5io.read_dataset(os.path.join(H5, "datasets", str(c["dataset"])))
    if raw is None:
        raw = c["out_raw"]
        forces = c["out_forces"] if int(c["out_tforces"]) else None
    h5io.write_fnetout(path, mode, raw, ds["offsets"], ns.n_global, ns.n_out - ns.n_global,
                       forces, ds["globaltargets"], ds["atomictargets"])
    return mode


def _drop_empty(t):
    return {k: v for k, v in t.items()
            if not (v["kind"] == "dataset" and 0 in v["shape"] and k.endswith("atomictargets"))}


@needs_h5py
@pytest.mark.parametrize("case", FNETOUT_CASES)
def test_written_fnetout_has_the_reference_tree(case, tmp_path):
    p = tmp_path / "fnetout.hdf5"
    _write_case_fnetout(case, p)
    assert tree(p) == _drop_empty(tree(os.path.join(H5, case, "_fnetout.hdf5")))


def test_errors(tmp_path):
    with pytest.raises(fa.FnetError, match="Invalid program running mode selected."):
        h5io.write_fnetout(tmp_path / "o.hdf5", "train", np.zeros((1, 1)), [0, 1], 1, 0)
    with pytest.raises(fa.FnetError, match="Validation mode only valid in combination with "
                                           "target data."):
        h5io.write_fnetout(tmp_path / "o.hdf5", "validate", np.zeros((1, 1)), [0, 1], 0, 0)
    with pytest.raises(fa.FnetError, match="cannot open HDF5 file"):
        h5io.read_netstat(tmp_path / "missing.hdf5")
    case = SD_CASES[0]
    src = os.path.join(H5, case, "fortnet.hdf5")
    bad = tmp_path / "bad.hdf5"
    shutil.copyfile(src, bad)
    with h5io.H5File(bad, "r+") as f:
        f.set_attr_string("netstat/mapping", "type", "soap")
    with pytest.raises(fa.FnetError, match="Invalid mapping type soap"):
        h5io.read_netstat(bad)
    shutil.copyfile(src, bad)
    with h5io.H5File(bad, "r+") as f:
        f.set_attr_int("netstat/mapping/function1", "tangular",
                       int(f.attr("netstat/mapping/function1", "tradial")))
    with pytest.raises(fa.FnetError, match="Conflicting entries tRadial and tAngular"):
        h5io.read_netstat(bad)
    shutil.copyfile(src, bad)
    with h5io.H5File(bad, "r+") as f:
        f.set_attr_int("netstat/bpnn", "nglobaltargets", 0)
        f.set_attr_int("netstat/bpnn", "natomictargets", 0)
    with pytest.raises(fa.FnetError, match="Total number of targets is zero"):
        h5io.read_netstat(bad)


def test_file_primitives(tmp_path):
    p = tmp_path / "x.hdf5"
    with h5io.H5File(p, "w") as f:
        f.create_group("a")
        f.create_group("a/b")
        f.write_int("a/i", np.arange(6, dtype=np.int32).reshape(2, 3))
        f.write_double("a/b/d", np.linspace(0, 1, 5))
        f.set_attr_string("a", "s", "text")
        f.set_attr_int("a/b", "n", [1, 2, 3])
        f.set_attr_double(".", "x", 0.25)
        assert f.list("a") == ["b", "i"]
        assert f.list("a/b") == ["d"]
        assert f.exists("a/b/d") and not f.exists("a/c")
    with h5io.H5File(p) as f:
        assert f.attr("a", "s") == "text"
        assert list(f.attr("a/b", "n")) == [1, 2, 3]
        assert f.attr(".", "x") == 0.25
        assert f.attr("a", "missing", default=7) == 7
        assert f.info("a/i") == ((2, 3), h5io.INT)
        assert _equal_arrays(f.dataset("a/i"), np.arange(6).reshape(2, 3).astype(np.int32))
        assert _equal_arrays(f.dataset("a/b/d"), np.linspace(0, 1, 5))
        assert _equal_arrays(f.dataset("a/i", np.float64), np.arange(6.0).reshape(2, 3))
    with h5io.H5File(p, "r+") as f:
        f.delete("a/i")
        assert f.list("a") == ["b"]


FORCE_CASES = [c for c in FNETOUT_CASES if c.startswith("analysis__derivatives__analytical")]


@pytest.mark.gpu
@pytest.mark.parametrize("case", FORCE_CASES + ["validation__multispecies__bothTargets",
                                                "prediction__multispecies__globalTargets"
                                                "__acsffeatures__speciesResolved"])
def test_predict_files_match_reference_fnetout(case, tmp_path):
    """The reference's prediction / validation run on the device, files in and out: the written
    fnetout holds the reference's predictions and analytic forces at its tolerance."""
    print("case", case, flush=True)
    c, _ = load_case(case)
    out = tmp_path / "fnetout.hdf5"
    ref = h5io.read_fnetout(os.path.join(H5, case, "_fnetout.hdf5"))
    h5io.predict(os.path.join(H5, "datasets", str(c["dataset"])),
                 os.path.join(H5, case, "fortnet.hdf5"), out, mode=ref["mode"],
                 forces=bool(ref["tforces"]))
    got = h5io.read_fnetout(out)
    assert got["mode"] == ref["mode"]
    assert np.array_equal(got["offsets"], ref["offsets"])
    for k in ("raw", "glob", "forces"):
        if k not in ref:
            continue
        want = ref[k]
        # element-wise, as testwithworkdir.py:700-714 compares
        assert np.allclose(got[k], want, atol=ATOL, rtol=RTOL), (k, np.abs(got[k] - want).max())
