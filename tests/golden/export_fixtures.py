"""Convert the reference's HDF5 regression fixtures into small ``.npz`` files.

This script is run once, in the development container, with an interpreter that
has ``h5py`` (``/opt/conda/bin/python3.9 tests/golden/export_fixtures.py``).  It
reads data only -- the datasets, netstat and fnetout HDF5 files held under
``test/prog/fortnet`` of the reference -- and writes plain arrays, so nothing of
the reference is executed or copied as source.  The GPU tests never read the
reference; they only see the ``.npz`` files written here.

Layout conventions (Fortran writes column-major arrays, h5py shows them
transposed):

* dataset ``coordinates`` h5py ``(nAtom, 3)``  == Fortran ``coords(3, nAtom)``;
* netstat ``layer{l}/weights`` h5py ``(d_{l+1}, d_l)`` == Fortran ``ww(d_l, d_{l+1})``
  (``prog/fortnet/lib_nn/bpnn.F90:1113-1116``), ``layer{l}/bias`` holds the bias
  of layer ``l+1``;
* fnetout ``forces`` h5py ``(nAtom, 3*nOut)`` == Fortran ``forces(3*nOut, nAtom)``
  (``prog/fortnet/lib_analysis/forces.F90:372-375``).

The ``.npz`` files store everything row-major with the atom index first, which
is the same memory order as the Fortran arrays.
"""
import os
import re
import sys

import h5py
import numpy as np

REF = "/root/reference/test/prog/fortnet"
OUT = os.path.dirname(os.path.abspath(__file__))

FUNC_TYPES = {"g1": 1, "g2": 2, "g3": 3, "g4": 4, "g5": 5}
ACTIVATIONS = ["tanh", "sigmoid", "relu", "lrelu", "softplus", "bent", "atan", "gaussian",
               "heaviside", "linear"]


def _attr(obj, key, default=None):
    if key not in obj.attrs:
        return default
    val = obj.attrs[key]
    if isinstance(val, bytes):
        return val.decode()
    if isinstance(val, np.ndarray) and val.size == 1:
        return val.reshape(-1)[0]
    return val


def export_dataset(path):
    """Read an fnetdata HDF5 file (``prog/fortnet/lib_io/fnetdata.F90:994-1216``)."""
    with h5py.File(path, "r") as f:
        ds = f["fnetdata/dataset"]
        ndp = int(_attr(ds, "ndatapoints"))
        next = int(_attr(ds, "nextfeatures", 0))
        withstruct = int(_attr(ds, "withstructures", 1))
        tr = f["fnetdata/dataset/training"]
        nglob = int(_attr(tr, "nglobaltargets", 0))
        natomic = int(_attr(tr, "natomictargets", 0))
        atnums = np.asarray(ds["atomicnumbers"][()], dtype=np.int32).reshape(-1)
        natoms, basis, periodic, frac, coords, gsp, atnum = [], [], [], [], [], [], []
        ext, aw, gt, at, weights = [], [], [], [], []
        for i in range(1, ndp + 1):
            dp = ds["datapoint%d" % i]
            weights.append(int(_attr(dp, "weight", 1)))
            geo = dp["geometry"]
            xyz = np.asarray(geo["coordinates"][()], dtype=np.float64)
            n = xyz.shape[0]
            natoms.append(n)
            per = int(_attr(geo, "periodic", 0))
            fr = int(_attr(geo, "fractional", 0))
            periodic.append(per)
            frac.append(fr)
            if per:
                basis.append(np.asarray(geo["basis"][()], dtype=np.float64))
            else:
                basis.append(np.zeros((3, 3)))
            coords.append(xyz)
            gsp.append(np.asarray(geo["localattoglobalsp"][()], dtype=np.int32).reshape(-1))
            atnum.append(np.asarray(geo["localattoatnum"][()], dtype=np.int32).reshape(-1))
            if next > 0:
                ext.append(np.asarray(dp["extfeatures"][()], dtype=np.float64).reshape(n, next))
            else:
                ext.append(np.zeros((n, 0)))
            if "atomicweights" in dp:
                aw.append(np.asarray(dp["atomicweights"][()], dtype=np.float64).reshape(n))
            else:
                aw.append(np.ones(n))
            if nglob > 0:
                gt.append(np.asarray(dp["globaltargets"][()], dtype=np.float64).reshape(nglob))
            else:
                gt.append(np.zeros(0))
            if natomic > 0:
                at.append(np.asarray(dp["atomictargets"][()], dtype=np.float64).reshape(n, natomic))
            else:
                at.append(np.zeros((n, 0)))
    off = np.concatenate([[0], np.cumsum(natoms)]).astype(np.int64)
    return dict(
        ndatapoints=ndp, natoms=np.array(natoms, np.int32), offsets=off,
        basis=np.stack(basis), periodic=np.array(periodic, np.int32),
        fractional=np.array(frac, np.int32), coords=np.concatenate(coords),
        globalsp=np.concatenate(gsp), atnum=np.concatenate(atnum),
        atomicnumbers=atnums, nextfeatures=next, ext=np.concatenate(ext),
        atomicweights=np.concatenate(aw), nglobaltargets=nglob, natomictargets=natomic,
        globaltargets=np.stack(gt) if nglob > 0 else np.zeros((ndp, 0)),
        atomictargets=np.concatenate(at), weights=np.array(weights, np.int32),
        withstructures=withstruct)


def export_netstat(path):
    """Read a netstat file: ACSF mapping (``acsf.F90:1908-2135``) + BPNN (``bpnn.F90:1141-1288``)."""
    out = {}
    with h5py.File(path, "r") as f:
        ns = f["netstat"]
        if "mapping" in ns:
            mp = ns["mapping"]
            nf = int(_attr(mp, "nfunctions"))
            cols = {k: [] for k in ("type", "rcut", "eta", "rs", "kappa", "lam", "xi", "atomid",
                                    "z1", "z2")}
            for i in range(1, nf + 1):
                g = mp["function%d" % i]
                cols["type"].append(FUNC_TYPES[_attr(g, "type").strip().lower()])
                cols["rcut"].append(float(_attr(g, "cutoff")))
                cols["eta"].append(float(_attr(g, "eta", 0.0)))
                cols["rs"].append(float(_attr(g, "rs", 0.0)))
                cols["kappa"].append(float(_attr(g, "kappa", 0.0)))
                cols["lam"].append(float(_attr(g, "lambda", 0.0)))
                cols["xi"].append(float(_attr(g, "xi", 0.0)))
                cols["atomid"].append(int(_attr(g, "atomid", 0)))
                z = np.asarray(g.attrs["atomicnumbers"]).reshape(-1)
                cols["z1"].append(int(z[0]))
                cols["z2"].append(int(z[1]))
            for k, v in cols.items():
                dt = np.int32 if k in ("type", "atomid", "z1", "z2") else np.float64
                out["g_" + k] = np.array(v, dtype=dt)
            if "preconditioning" in mp:
                pc = mp["preconditioning"]
                out["zmean"] = np.asarray(pc["means"][()], np.float64)
                out["zstd"] = np.asarray(pc["variances"][()], np.float64)
        if "external" in ns:
            out["ext_indices"] = np.asarray(ns["external/indices"][()], np.int32).reshape(-1)
        bp = ns["bpnn"]
        out["nglobaltargets"] = int(_attr(bp, "nglobaltargets"))
        out["natomictargets"] = int(_attr(bp, "natomictargets"))
        atn = np.asarray(bp["atomicnumbers"][()], np.int32).reshape(-1)
        out["net_atomicnumbers"] = atn
        acts = []
        params = []
        topo = None
        for z in atn:
            sub = [k for k in bp.keys() if k.endswith("-subnetwork")
                   and int(_attr(bp[k], "element")) == int(z)]
            assert len(sub) == 1, (path, z, list(bp.keys()))
            g = bp[sub[0]]
            acts.append(ACTIVATIONS.index(_attr(g, "activation").strip().lower()))
            t = np.asarray(g["topology"][()], np.int32).reshape(-1)
            if topo is None:
                topo = t
            assert np.array_equal(topo, t)
            ww = []
            bb = [np.zeros(int(t[0]))]
            for il in range(1, len(t)):
                w = np.asarray(g["layer%d/weights" % il][()], np.float64).reshape(t[il], t[il - 1]).T
                ww.append(w)
                bb.append(np.asarray(g["layer%d/bias" % il][()], np.float64).reshape(t[il]))
            # serialised order of network.F90:397-427: per layer ww(:, ii) columns, including
            # the dummy output layer ww(d_n, 1) (always zero), then all biases
            vec = []
            for w in ww:
                vec.append(w.T.reshape(-1))  # column-major of ww(d_l, d_{l+1})
            vec.append(np.zeros(int(t[-1])))
            vec.extend(bb)
            params.append(np.concatenate(vec))
        out["topology"] = topo
        out["activation"] = np.array(acts, np.int32)
        out["params"] = np.stack(params)
    return out


def export_fnetout(path):
    """Read the expected outputs (``prog/fortnet/lib_io/fnetout.F90:33-158``)."""
    with h5py.File(path, "r") as f:
        o = f["fnetout/output"]
        ndp = int(_attr(o, "ndatapoints"))
        tforces = int(_attr(o, "tforces", 0))
        raw, glob, forces = [], [], []
        for i in range(1, ndp + 1):
            d = o["datapoint%d" % i]
            r = np.asarray(d["rawpredictions"][()], np.float64)
            raw.append(r)
            if "globalpredictions" in d:
                glob.append(np.asarray(d["globalpredictions"][()], np.float64).reshape(-1))
            if tforces:
                forces.append(np.asarray(d["forces"][()], np.float64))
    res = {"raw": np.concatenate(raw), "tforces": tforces}
    if glob:
        res["glob"] = np.stack(glob)
    if forces:
        res["forces"] = np.concatenate(forces)
    return res


def parse_hsd(path):
    """Pick the few input keys the fixtures depend on (``initprogram.F90`` defaults)."""
    txt = open(path).read()
    info = {}
    m = re.search(r"Dataset\s*=\s*'?([^'\n]+?)'?\s*\n", txt)
    info["dataset"] = os.path.basename(m.group(1).strip()) if m else ""
    m = re.search(r"Mode\s*=\s*'?(\w+)'?", txt)
    info["mode"] = m.group(1).lower() if m else ""
    m = re.search(r"Training\s*=\s*(\w+)", txt)
    info["optimizer"] = m.group(1).lower() if m else ""
    m = re.search(r"LearningRate\s*=\s*([0-9.eE+-]+)", txt)
    info["learningrate"] = float(m.group(1)) if m else 0.0
    m = re.search(r"MaxDisplacement\s*=\s*([0-9.eE+-]+)", txt)
    info["maxdisp"] = float(m.group(1)) if m else 0.0
    m = re.search(r"Loss\s*=\s*'?(\w+)'?", txt)
    info["loss"] = m.group(1).lower() if m else "mse"
    m = re.search(r"Forces\s*=\s*(\w+)", txt)
    info["forces"] = m.group(1).lower() if m else ""
    m = re.search(r"Delta\s*=\s*([0-9.eE+-]+)", txt)
    info["delta"] = float(m.group(1)) if m else 0.0
    m = re.search(r"Shuffle\s*=\s*(\w+)", txt)
    info["shuffle"] = m.group(1).lower() if m else "no"
    m = re.search(r"Regularization\s*=\s*(\w+)", txt)
    info["regularization"] = 1 if m else 0
    info["regu_type"] = m.group(1).lower() if m else ""
    m = re.search(r"Strength\s*=\s*([0-9.eE+-]+)", txt)
    info["regu_strength"] = float(m.group(1)) if m else 0.0
    m = re.search(r"Alpha\s*=\s*([0-9.eE+-]+)", txt)
    info["regu_alpha"] = float(m.group(1)) if m else 0.0
    m = re.search(r"RandomSeed\s*=\s*([0-9]+)", txt)
    info["seed"] = int(m.group(1)) if m else 0
    m = re.search(r"Threshold\s*=\s*([0-9.eE+-]+)", txt)
    info["threshold"] = float(m.group(1)) if m else 0.0
    m = re.search(r"MinDisplacement\s*=\s*([0-9.eE+-]+)", txt)
    info["mindisp"] = float(m.group(1)) if m else 1e-6
    m = re.search(r"NIterations\s*=\s*([0-9]+)", txt)
    info["niterations"] = int(m.group(1)) if m else 0
    m = re.search(r"Memory\s*=\s*([0-9]+)", txt)
    info["memory"] = int(m.group(1)) if m else 1000
    m = re.search(r"Linemin\s*=\s*(\w+)", txt)
    info["linemin"] = 0 if (m and m.group(1).lower() == "no") else 1
    m = re.search(r"ReadNetStats\s*=\s*(\w+)", txt)
    info["readnetstats"] = 1 if (m and m.group(1).lower() == "yes") else 0
    return info


def main():
    os.makedirs(os.path.join(OUT, "datasets"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "cases"), exist_ok=True)
    for fn in sorted(os.listdir(os.path.join(REF, "datasets"))):
        d = export_dataset(os.path.join(REF, "datasets", fn))
        np.savez_compressed(os.path.join(OUT, "datasets", fn.replace(".hdf5", ".npz")), **d)
    n = 0
    roots = ["analysis", "prediction", "validation", "restart", "input", "training"]
    for root in roots:
        for dirpath, _, files in sorted(os.walk(os.path.join(REF, root))):
            if "fortnet_in.hsd" not in files:
                continue
            rel = os.path.relpath(dirpath, REF)
            info = parse_hsd(os.path.join(dirpath, "fortnet_in.hsd"))
            # training from scratch (RANLUX-initialised weights) has no input netstat
            scratch = info["mode"] == "train" and not info["readnetstats"]
            if "fortnet.hdf5" not in files and not scratch:
                continue
            if scratch and "_fortnet.hdf5" not in files:
                continue
            case = {"name": rel, "dataset": info["dataset"], "mode": info["mode"],
                    "optimizer": info["optimizer"], "learningrate": info["learningrate"],
                    "maxdisp": info["maxdisp"],
                    "loss": info["loss"], "forces_kind": info["forces"], "delta": info["delta"],
                    "shuffle": info["shuffle"], "regularization": info["regularization"],
                    "regu_type": info["regu_type"], "regu_strength": info["regu_strength"],
                    "regu_alpha": info["regu_alpha"], "seed": info["seed"],
                    "threshold": info["threshold"], "mindisp": info["mindisp"],
                    "niterations": info["niterations"], "memory": info["memory"],
                    "linemin": info["linemin"], "scratch": int(scratch)}
            if "fortnet.hdf5" in files:
                for k, v in export_netstat(os.path.join(dirpath, "fortnet.hdf5")).items():
                    case["in_" + k] = v
            if "_fnetout.hdf5" in files:
                for k, v in export_fnetout(os.path.join(dirpath, "_fnetout.hdf5")).items():
                    case["out_" + k] = v
            if info["mode"] == "train" and "_fortnet.hdf5" in files:
                for k, v in export_netstat(os.path.join(dirpath, "_fortnet.hdf5")).items():
                    case["after_" + k] = v
            name = rel.replace("/", "__")
            np.savez_compressed(os.path.join(OUT, "cases", name + ".npz"), **case)
            n += 1
    print("exported", n, "cases")


if __name__ == "__main__":
    sys.exit(main())
