"""Dump the tree of an HDF5 file as JSON: every group and dataset with its attributes, and every
dataset's type, shape and values.  Uses h5py, so it runs in an interpreter that has it (the
tests call it as an independent reader of the files the HDF5 layer writes)."""
import json
import sys

import h5py
import numpy as np


def _value(v):
    if isinstance(v, bytes):
        return ["str", v.decode().rstrip(" \0")]
    if isinstance(v, str):
        return ["str", v.rstrip(" \0")]
    a = np.asarray(v)
    if a.dtype.kind == "S":
        return ["str", a.reshape(-1)[0].decode().rstrip(" \0")]
    if a.dtype.kind in "UO":
        x = a.reshape(-1)[0]
        return ["str", (x.decode() if isinstance(x, bytes) else str(x)).rstrip(" \0")]
    return [a.dtype.kind + str(a.dtype.itemsize), list(a.shape), a.reshape(-1).tolist()]


def dump(path):
    out = {}

    def attrs(obj):
        return {k: _value(obj.attrs[k]) for k in obj.attrs.keys()}

    def visit(name, obj):
        if isinstance(obj, h5py.Group):
            out[name] = {"kind": "group", "attrs": attrs(obj)}
        else:
            a = obj[()]
            out[name] = {"kind": "dataset", "dtype": a.dtype.kind + str(a.dtype.itemsize),
                         "shape": list(a.shape), "values": a.reshape(-1).tolist(),
                         "attrs": attrs(obj)}

    with h5py.File(path, "r") as f:
        out["/"] = {"kind": "group", "attrs": attrs(f)}
        f.visititems(visit)
    return out


if __name__ == "__main__":
    json.dump(dump(sys.argv[1]), sys.stdout)
