"""The reference's HDF5 files: fnetdata datasets, netstats and fnetout.

Built on the C-ABI of ``libfnet_h5.so`` (``include/fnet_h5.h``), which wraps the HDF5 C library
the way the reference's ``lib_hdf5fx/hdf5fx.F90`` wraps the Fortran one.  The schema follows the
reference's readers and writers:

* :func:`read_dataset` -- ``readHdfDataset`` (``lib_io/fnetdata.F90:994-1216``): group
  ``fnetdata/dataset`` with ``datapoint<i>/geometry`` per structure;
* :func:`read_netstat` / :func:`write_netstat` -- ``netstat.F90:37-453`` (file, ``bpnn`` header,
  ``external``), ``TBpnn_toFile`` / ``TBpnn_fromFile`` (``lib_nn/bpnn.F90:1062-1288``) and
  ``TAcsf_toFile`` / ``TAcsf_fromFile`` (``lib_descriptors/acsf.F90:1759-2135``);
* :func:`write_fnetout` / :func:`read_fnetout` -- ``writeFnetout`` (``lib_io/fnetout.F90:33-158``);
* :func:`predict` -- a dataset and a netstat in, an fnetout out, on the device: the reference's
  prediction / validation run for the analytic-force path.

Arrays are returned and taken in C order with the dimensions HDF5 stores, which for the
reference's Fortran arrays are the Fortran dimensions reversed: coordinates ``(nAtom, 3)``,
lattice ``(3, 3)`` with the vectors as rows, raw predictions ``(nAtom, nOut)``, forces
``(nAtom, 3 nOut)``, layer weights ``(d_{l+1}, d_l)``.  The dictionaries of :func:`read_dataset`
and :func:`read_netstat` have the keys of the fixture conversion, so :func:`dataset.geometries`
and :meth:`Netstat.from_arrays` take them unchanged.
"""
import ctypes
import os

import numpy as np

from ._lib import FnetError
from .core import ACTIVATIONS, Context, Netstat, nparams
from . import dataset as _dataset

HERE = os.path.dirname(os.path.abspath(__file__))
H5_LIB_PATH = os.path.join(HERE, "libfnet_h5.so")

INT, FLOAT, STRING, OTHER = 0, 1, 2, 3
MAX_RANK = 8

# Element symbols 1..118 as the reference names its sub-networks (lower case,
# lib_dftbp/constants.F90:156-173): "<symbol>-subnetwork".
_SYMBOLS = (
    "h he li be b c n o f ne na mg al si p s cl ar k ca sc ti v cr mn fe co ni cu zn ga ge as se "
    "br kr rb sr y zr nb mo tc ru rh pd ag cd in sn sb te i xe cs ba la ce pr nd pm sm eu gd tb dy "
    "ho er tm yb lu hf ta w re os ir pt au hg tl pb bi po at rn fr ra ac th pa u np pu am cm bk cf "
    "es fm md no lr rf db sg bh hs mt ds rg cn nh fl mc lv ts og").split()
assert len(_SYMBOLS) == 118


def element_symbol(z):
    z = int(z)
    if not 1 <= z <= 118:
        raise FnetError("no element symbol for atomic number %d" % z)
    return _SYMBOLS[z - 1]


FUNC_TYPES = {"g1": 1, "g2": 2, "g3": 3, "g4": 4, "g5": 5}

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(H5_LIB_PATH):
        raise FnetError("libfnet_h5.so is not built (run `python -m fortnet_amd.build`; needs "
                        "the HDF5 C and high-level libraries)")
    try:
        L = ctypes.CDLL(H5_LIB_PATH)
    except OSError as e:
        raise FnetError("cannot load libfnet_h5.so (HDF5 libraries missing?): %s" % e)
    c_ll, c_int, c_str, c_p = ctypes.c_longlong, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p
    sig = {
        "fnet_h5_last_error": ([], c_str),
        "fnet_h5_create": ([c_str], c_ll),
        "fnet_h5_open": ([c_str, c_int], c_ll),
        "fnet_h5_close": ([c_ll], c_int),
        "fnet_h5_group_create": ([c_ll, c_str], c_int),
        "fnet_h5_exists": ([c_ll, c_str], c_int),
        "fnet_h5_delete": ([c_ll, c_str], c_int),
        "fnet_h5_list": ([c_ll, c_str, c_p, c_ll], c_int),
        "fnet_h5_attr_exists": ([c_ll, c_str, c_str], c_int),
        "fnet_h5_attr_info": ([c_ll, c_str, c_str, c_p, c_p, c_p], c_int),
        "fnet_h5_set_attr_int": ([c_ll, c_str, c_str, c_p, c_ll], c_int),
        "fnet_h5_set_attr_double": ([c_ll, c_str, c_str, c_p, c_ll], c_int),
        "fnet_h5_set_attr_string": ([c_ll, c_str, c_str, c_str], c_int),
        "fnet_h5_get_attr_int": ([c_ll, c_str, c_str, c_p, c_ll], c_int),
        "fnet_h5_get_attr_double": ([c_ll, c_str, c_str, c_p, c_ll], c_int),
        "fnet_h5_get_attr_string": ([c_ll, c_str, c_str, c_p, c_ll], c_int),
        "fnet_h5_make_dataset_int": ([c_ll, c_str, c_int, c_p, c_p], c_int),
        "fnet_h5_make_dataset_double": ([c_ll, c_str, c_int, c_p, c_p], c_int),
        "fnet_h5_dataset_info": ([c_ll, c_str, c_p, c_p, c_p], c_int),
        "fnet_h5_read_dataset_int": ([c_ll, c_str, c_p, c_ll], c_int),
        "fnet_h5_read_dataset_double": ([c_ll, c_str, c_p, c_ll], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _b(s):
    return s.encode()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class H5File:
    """An HDF5 file: ``mode`` "r" (read), "r+" (read and write) or "w" (create, truncate)."""

    def __init__(self, path, mode="r"):
        self._L = _load()
        self.path = str(path)
        if mode == "w":
            h = self._L.fnet_h5_create(_b(self.path))
        elif mode in ("r", "r+"):
            h = self._L.fnet_h5_open(_b(self.path), 1 if mode == "r+" else 0)
        else:
            raise ValueError("mode must be 'r', 'r+' or 'w'")
        self._check(h)
        self.h = h

    def _check(self, rc):
        if rc < 0:
            raise FnetError(self._L.fnet_h5_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None) is not None:
            h, self.h = self.h, None
            self._check(self._L.fnet_h5_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- structure ----
    def exists(self, path):
        return bool(self._check(self._L.fnet_h5_exists(self.h, _b(path))))

    def delete(self, path):
        self._check(self._L.fnet_h5_delete(self.h, _b(path)))

    def create_group(self, path):
        self._check(self._L.fnet_h5_group_create(self.h, _b(path)))

    def list(self, path="."):
        need = self._L.fnet_h5_list(self.h, _b(path), None, 0)
        if need >= 0:  # empty group
            return []
        if need == -1:
            self._check(need)
        buf = ctypes.create_string_buffer(-need - 1)
        n = self._check(self._L.fnet_h5_list(self.h, _b(path), buf, -need - 1))
        names = buf.raw.split(b"\0")[:n]
        return [x.decode() for x in names]

    # ---- attributes ----
    def attr_exists(self, obj, name):
        return bool(self._check(self._L.fnet_h5_attr_exists(self.h, _b(obj), _b(name))))

    def attr(self, obj, name, default=None):
        """Value of an attribute: str for strings, else a 1-D int32 / float64 array (a scalar
        for a single value).  ``default`` when the attribute is absent."""
        if default is not None and not self.attr_exists(obj, name):
            return default
        n, cls, size = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_longlong()
        self._check(self._L.fnet_h5_attr_info(self.h, _b(obj), _b(name), ctypes.byref(n),
                                              ctypes.byref(cls), ctypes.byref(size)))
        if cls.value == STRING:
            cap = max(int(size.value), 0) + 1 if size.value >= 0 else 4096
            buf = ctypes.create_string_buffer(cap)
            self._check(self._L.fnet_h5_get_attr_string(self.h, _b(obj), _b(name), buf, cap))
            return buf.value.decode()
        if cls.value == INT:
            v = np.zeros(n.value, np.int32)
            self._check(self._L.fnet_h5_get_attr_int(self.h, _b(obj), _b(name), _ptr(v),
                                                     n.value))
        elif cls.value == FLOAT:
            v = np.zeros(n.value, np.float64)
            self._check(self._L.fnet_h5_get_attr_double(self.h, _b(obj), _b(name), _ptr(v),
                                                        n.value))
        else:
            raise FnetError("attribute '%s' of '%s' has an unsupported type" % (name, obj))
        return v[0] if v.size == 1 else v

    def set_attr_int(self, obj, name, values):
        v = np.ascontiguousarray(np.atleast_1d(values), np.int32)
        self._check(self._L.fnet_h5_set_attr_int(self.h, _b(obj), _b(name), _ptr(v), v.size))

    def set_attr_double(self, obj, name, values):
        v = np.ascontiguousarray(np.atleast_1d(values), np.float64)
        self._check(self._L.fnet_h5_set_attr_double(self.h, _b(obj), _b(name), _ptr(v),
                                                     v.size))

    def set_attr_string(self, obj, name, text):
        self._check(self._L.fnet_h5_set_attr_string(self.h, _b(obj), _b(name), _b(text)))

    # ---- datasets ----
    def info(self, path):
        rank, cls = ctypes.c_int(), ctypes.c_int()
        dims = np.zeros(MAX_RANK, np.int64)
        self._check(self._L.fnet_h5_dataset_info(self.h, _b(path), ctypes.byref(rank),
                                                 _ptr(dims), ctypes.byref(cls)))
        return tuple(int(x) for x in dims[:rank.value]), cls.value

    def dataset(self, path, dtype=None):
        """A dataset as an array of its stored shape: int32 for integer data, float64 for
        floating point (or ``dtype``, which converts on reading as HDF5 does)."""
        shape, cls = self.info(path)
        if dtype is None:
            dtype = np.int32 if cls == INT else np.float64
        dtype = np.dtype(dtype)
        out = np.zeros(shape, dtype)
        n = out.size
        if dtype == np.int32:
            self._check(self._L.fnet_h5_read_dataset_int(self.h, _b(path), _ptr(out), n))
        elif dtype == np.float64:
            self._check(self._L.fnet_h5_read_dataset_double(self.h, _b(path), _ptr(out), n))
        else:
            raise ValueError("int32 or float64 only")
        return out

    def write_int(self, path, array):
        a = np.ascontiguousarray(array, np.int32)
        if a.ndim == 0:
            a = a.reshape(1)
        dims = np.array(a.shape, np.int64)
        self._check(self._L.fnet_h5_make_dataset_int(self.h, _b(path), a.ndim, _ptr(dims),
                                                     _ptr(a)))

    def write_double(self, path, array):
        a = np.ascontiguousarray(array, np.float64)
        if a.ndim == 0:
            a = a.reshape(1)
        dims = np.array(a.shape, np.int64)
        self._check(self._L.fnet_h5_make_dataset_double(self.h, _b(path), a.ndim, _ptr(dims),
                                                        _ptr(a)))


# ---------------------------------------------------------------------------------------------
# fnetdata datasets

def read_dataset(path):
    """``readHdfDataset`` (fnetdata.F90:994-1216): every datapoint of ``fnetdata/dataset``, atoms
    concatenated; the keys are those of the fixture conversion (``offsets`` per structure,
    ``coords`` as stored, fractional or Cartesian, ``basis`` with the vectors as rows)."""
    with H5File(path) as f:
        root = "fnetdata/dataset"
        if not f.exists(root):
            raise FnetError("'%s' has no fnetdata/dataset group" % path)
        ndp = int(f.attr(root, "ndatapoints"))
        n_ext = int(f.attr(root, "nextfeatures", 0))
        withstruct = int(f.attr(root, "withstructures", 1))
        tr = root + "/training"
        nglob = int(f.attr(tr, "nglobaltargets", 0))
        natomic = int(f.attr(tr, "natomictargets", 0))
        atnums = f.dataset(root + "/atomicnumbers", np.int32).reshape(-1)
        natoms, basis, periodic, frac, coords, gsp, atnum = [], [], [], [], [], [], []
        ext, aw, gt, at, weights = [], [], [], [], []
        for i in range(1, ndp + 1):
            dp = "%s/datapoint%d" % (root, i)
            weights.append(int(f.attr(dp, "weight", 1)))
            geo = dp + "/geometry"
            xyz = f.dataset(geo + "/coordinates", np.float64).reshape(-1, 3)
            n = xyz.shape[0]
            natoms.append(n)
            per = int(f.attr(geo, "periodic", 0))
            fr = int(f.attr(geo, "fractional", 0))
            periodic.append(per)
            frac.append(fr)
            if per:
                lat = f.dataset(geo + "/basis", np.float64).reshape(3, 3)
                if abs(np.linalg.det(lat)) < 1e-12:  # fnetdata.F90:1153-1156
                    raise FnetError("Dependent lattice vectors in datapoint%d." % i)
                basis.append(lat)
            else:
                basis.append(np.zeros((3, 3)))
            coords.append(xyz)
            gsp.append(f.dataset(geo + "/localattoglobalsp", np.int32).reshape(-1))
            atnum.append(f.dataset(geo + "/localattoatnum", np.int32).reshape(-1))
            ext.append(f.dataset(dp + "/extfeatures", np.float64).reshape(n, n_ext) if n_ext > 0
                       else np.zeros((n, 0)))
            aw.append(f.dataset(dp + "/atomicweights", np.float64).reshape(n)
                      if f.exists(dp + "/atomicweights") else np.ones(n))
            gt.append(f.dataset(dp + "/globaltargets", np.float64).reshape(nglob) if nglob > 0
                      else np.zeros(0))
            at.append(f.dataset(dp + "/atomictargets", np.float64).reshape(n, natomic)
                      if natomic > 0 else np.zeros((n, 0)))
    off = np.concatenate([[0], np.cumsum(natoms)]).astype(np.int64)
    return dict(
        ndatapoints=ndp, natoms=np.array(natoms, np.int32), offsets=off,
        basis=np.stack(basis), periodic=np.array(periodic, np.int32),
        fractional=np.array(frac, np.int32), coords=np.concatenate(coords),
        globalsp=np.concatenate(gsp), atnum=np.concatenate(atnum),
        atomicnumbers=atnums, nextfeatures=n_ext, ext=np.concatenate(ext),
        atomicweights=np.concatenate(aw), nglobaltargets=nglob, natomictargets=natomic,
        globaltargets=np.stack(gt) if nglob > 0 else np.zeros((ndp, 0)),
        atomictargets=np.concatenate(at), weights=np.array(weights, np.int32),
        withstructures=withstruct)


# ---------------------------------------------------------------------------------------------
# netstat

def _layer_slices(dims):
    """Offsets of each layer's weights ww(d_l, d_{l+1}) and each bias bb(d_l) in the serialised
    parameter vector (network.F90:397-427: all weight matrices column by column, the dummy output
    matrix ww(d_n, 1), then every layer's biases)."""
    dims = [int(d) for d in dims]
    w, o = [], 0
    for l in range(len(dims) - 1):
        w.append((o, dims[l] * dims[l + 1]))
        o += dims[l] * dims[l + 1]
    o += dims[-1]
    b = []
    for d in dims:
        b.append((o, d))
        o += d
    assert o == nparams(dims)
    return w, b


def read_netstat(path):
    """Netstat arrays (ACSF mapping and z-score, external inputs, networks) with the keys of
    :meth:`Netstat.from_arrays`; the checks of ``TAcsf_fromFile`` (acsf.F90:1931-2064) and
    ``TBpnn_fromFile`` (bpnn.F90:1166-1231) raise :class:`FnetError` with the reference's
    messages."""
    out = {}
    with H5File(path) as f:
        if not f.exists("netstat"):
            raise FnetError("'%s' has no netstat group" % path)
        mp = "netstat/mapping"
        if f.exists(mp):
            kind = f.attr(mp, "type").strip().lower()
            if kind != "acsf":
                raise FnetError("Error while reading ACSF from netstat file. Invalid mapping "
                                "type %s." % kind)
            nf = int(f.attr(mp, "nfunctions"))
            if nf <= 0:
                raise FnetError("Error while reading ACSF from netstat file. Non-positive number "
                                "of ACSF functions obtained.")
            cols = {k: [] for k in ("type", "rcut", "eta", "rs", "kappa", "lam", "xi", "atomid",
                                    "z1", "z2")}
            for i in range(1, nf + 1):
                g = "%s/function%d" % (mp, i)
                t = f.attr(g, "type").strip().lower()
                if t not in FUNC_TYPES:
                    raise FnetError("Error while reading ACSF from netstat file. Invalid "
                                    "G-Function type %s." % t)
                atomid = int(f.attr(g, "atomid"))
                if atomid < 0:
                    raise FnetError("Error while reading ACSF from netstat file. Negative atomId "
                                    "obtained.")
                rc = float(f.attr(g, "cutoff"))
                if rc <= 0.0:
                    raise FnetError("Error while reading ACSF from netstat file. Invalid cutoff "
                                    "obtained.")
                z = np.atleast_1d(f.attr(g, "atomicnumbers")).astype(np.int32)
                if len(z) != 2 or z.min() < 0:
                    raise FnetError("Error while reading ACSF from netstat file. Invalid atomic "
                                    "numbers obtained.")
                trad, tang = int(f.attr(g, "tradial")), int(f.attr(g, "tangular"))
                if trad not in (0, 1):
                    raise FnetError("Error while reading ACSF from netstat file. Invalid tRadial "
                                    "entry.")
                if tang not in (0, 1):
                    raise FnetError("Error while reading ACSF from netstat file. Invalid "
                                    "tAngular entry.")
                if trad == tang:
                    raise FnetError("Error while reading ACSF from netstat file. Conflicting "
                                    "entries tRadial and tAngular.")
                ty = FUNC_TYPES[t]
                vals = dict(eta=0.0, rs=0.0, kappa=0.0, lam=0.0, xi=0.0)
                if ty == 2:
                    vals["eta"] = float(f.attr(g, "eta"))
                    vals["rs"] = float(f.attr(g, "rs"))
                elif ty == 3:
                    vals["kappa"] = float(f.attr(g, "kappa"))
                elif ty in (4, 5):
                    vals["xi"] = float(f.attr(g, "xi"))
                    vals["lam"] = float(f.attr(g, "lambda"))
                    vals["eta"] = float(f.attr(g, "eta"))
                cols["type"].append(ty)
                cols["rcut"].append(rc)
                cols["atomid"].append(atomid)
                cols["z1"].append(int(z[0]))
                cols["z2"].append(int(z[1]))
                for k, v in vals.items():
                    cols[k].append(v)
            for k, v in cols.items():
                dt = np.int32 if k in ("type", "atomid", "z1", "z2") else np.float64
                out["g_" + k] = np.array(v, dtype=dt)
            pc = mp + "/preconditioning"
            if f.exists(pc):
                kind = f.attr(pc, "type").strip().lower()
                if kind != "zscore":
                    raise FnetError("Error while reading ACSF from netstat file. Invalid "
                                    "preconditioning type %s." % kind)
                out["zmean"] = f.dataset(pc + "/means", np.float64).reshape(-1)
                out["zstd"] = f.dataset(pc + "/variances", np.float64).reshape(-1)
        if f.exists("netstat/external"):
            out["ext_indices"] = f.dataset("netstat/external/indices", np.int32).reshape(-1)
        bp = "netstat/bpnn"
        nglob = int(f.attr(bp, "nglobaltargets"))
        natomic = int(f.attr(bp, "natomictargets"))
        if nglob < 0:
            raise FnetError("Error while reading BPNN from netstat file. Negative number of "
                            "system-wide targets.")
        if natomic < 0:
            raise FnetError("Error while reading BPNN from netstat file. Negative number of "
                            "atomic targets.")
        if nglob + natomic == 0:
            raise FnetError("Error while reading BPNN from netstat file. Total number of targets "
                            "is zero")
        atn = f.dataset(bp + "/atomicnumbers", np.int32).reshape(-1)
        topo, act, params = None, None, []
        for z in atn:
            sub = "%s/%s-subnetwork" % (bp, element_symbol(z))
            t = f.dataset(sub + "/topology", np.int32).reshape(-1)
            if topo is not None and len(t) != len(topo):
                raise FnetError("Error while reading sub-network topology. Currently only "
                                "equally shaped networks are supported.")
            if topo is not None and not np.array_equal(t, topo):
                raise FnetError("Error while reading sub-network topology. Currently only equal "
                                "topologies are supported.")
            topo = t
            a = f.attr(sub, "activation").strip().lower()
            if act is not None and a != act:
                raise FnetError("Error while reading sub-network transfer functions. Currently "
                                "only networks with equal transfer are supported.")
            if a not in ACTIVATIONS:
                raise FnetError("unknown transfer function '%s'" % a)
            act = a
            wsl, bsl = _layer_slices(t)
            vec = np.zeros(nparams(t))
            for il in range(1, len(t)):
                lay = "%s/layer%d" % (sub, il)
                # weights (d_{l+1}, d_l) as stored = ww(d_l, d_{l+1}) column by column
                o, n = wsl[il - 1]
                vec[o:o + n] = f.dataset(lay + "/weights", np.float64).reshape(-1)
                o, n = bsl[il]
                vec[o:o + n] = f.dataset(lay + "/bias", np.float64).reshape(-1)
            params.append(vec)
    out["nglobaltargets"] = nglob
    out["natomictargets"] = natomic
    out["net_atomicnumbers"] = atn
    out["topology"] = topo
    out["activation"] = np.full(len(atn), ACTIVATIONS.index(act), np.int32)
    out["params"] = np.stack(params)
    return out


def load_netstat(path):
    """:class:`Netstat` of a netstat file."""
    return Netstat.from_arrays(read_netstat(path))


def write_netstat(path, ns, n_atomic=None):
    """Writes a :class:`Netstat` as the reference does: ``createNetstat`` (netstat.F90:37-60),
    ``TAcsf_toFile`` (acsf.F90:1759-1904), ``writeExtFeaturesConfig`` (netstat.F90:406-453),
    ``writeBpnnHeader`` (netstat.F90:312-402) and ``TBpnn_toFile`` (bpnn.F90:1062-1137).
    ``n_atomic`` defaults to the outputs beyond the system-wide ones."""
    nglob = int(ns.n_global)
    natomic = int(ns.n_out - nglob if n_atomic is None else n_atomic)
    with H5File(path, "w") as f:
        f.create_group("netstat")
        g = ns.g
        if len(g):
            mp = "netstat/mapping"
            f.create_group(mp)
            f.set_attr_string(mp, "type", "acsf")
            f.set_attr_int(mp, "nfunctions", len(g))
            for i in range(len(g)):
                fn = "%s/function%d" % (mp, i + 1)
                f.create_group(fn)
                ty = int(g.type[i])
                f.set_attr_string(fn, "type", "g%d" % ty)
                f.set_attr_int(fn, "atomid", int(g.atomid[i]))
                f.set_attr_double(fn, "cutoff", float(g.rcut[i]))
                f.set_attr_int(fn, "atomicnumbers", [int(g.z1[i]), int(g.z2[i])])
                f.set_attr_int(fn, "tradial", 1 if ty <= 3 else 0)
                f.set_attr_int(fn, "tangular", 0 if ty <= 3 else 1)
                if ty == 2:
                    f.set_attr_double(fn, "eta", float(g.eta[i]))
                    f.set_attr_double(fn, "rs", float(g.rs[i]))
                elif ty == 3:
                    f.set_attr_double(fn, "kappa", float(g.kappa[i]))
                elif ty in (4, 5):
                    f.set_attr_double(fn, "xi", float(g.xi[i]))
                    f.set_attr_double(fn, "lambda", float(g.lam[i]))
                    f.set_attr_double(fn, "eta", float(g.eta[i]))
            if ns.zmean is not None and ns.zstd is not None:
                pc = mp + "/preconditioning"
                f.create_group(pc)
                f.set_attr_string(pc, "type", "zscore")
                f.write_double(pc + "/means", ns.zmean)
                f.write_double(pc + "/variances", ns.zstd)
        if len(ns.ext_in):
            f.create_group("netstat/external")
            f.set_attr_int("netstat/external", "nextfeatures", len(ns.ext_in))
            f.write_int("netstat/external/indices", ns.ext_in)
        bp = "netstat/bpnn"
        f.create_group(bp)
        f.write_int(bp + "/atomicnumbers", ns.species_z)
        f.set_attr_int(bp, "nglobaltargets", nglob)
        f.set_attr_int(bp, "natomictargets", natomic)
        dims = [int(d) for d in ns.dims]
        wsl, bsl = _layer_slices(dims)
        for s, z in enumerate(ns.species_z):
            sub = "%s/%s-subnetwork" % (bp, element_symbol(z))
            f.create_group(sub)
            f.set_attr_int(sub, "element", int(z))
            f.set_attr_int(sub, "nlayer", len(dims))
            f.write_int(sub + "/topology", dims)
            f.set_attr_string(sub, "activation", ACTIVATIONS[int(ns.activation[s])])
            p = ns.params[s]
            for il in range(1, len(dims)):
                lay = "%s/layer%d" % (sub, il)
                f.create_group(lay)
                o, n = wsl[il - 1]
                f.write_double(lay + "/weights", p[o:o + n].reshape(dims[il], dims[il - 1]))
                o, n = bsl[il]
                f.write_double(lay + "/bias", p[o:o + n])


# ---------------------------------------------------------------------------------------------
# fnetout

def _sum_atoms(a):
    """Sum over atoms in index order, as the reference's sum(array(1:nG, :), dim=2)."""
    acc = np.zeros(a.shape[1])
    for row in a:
        acc = acc + row
    return acc


def write_fnetout(path, mode, raw, offsets, n_global, n_atomic, forces=None,
                  global_targets=None, atomic_targets=None):
    """``writeFnetout`` (fnetout.F90:33-158).  ``raw`` [nAtoms, nOut] and ``forces``
    [nAtoms, 3 nOut] over all structures (structure s owns rows offsets[s]:offsets[s+1]);
    ``global_targets`` [nStruct, n_global] and ``atomic_targets`` [nAtoms, n_atomic] for
    "validate"."""
    if mode not in ("validate", "predict"):
        raise FnetError("Invalid program running mode selected.")
    n_global, n_atomic = int(n_global), int(n_atomic)
    if mode == "validate" and n_global + n_atomic == 0:
        raise FnetError("Validation mode only valid in combination with target data.")
    raw = np.asarray(raw, np.float64)
    off = np.asarray(offsets, np.int64)
    ndp = len(off) - 1
    with H5File(path, "w") as f:
        f.create_group("fnetout")
        f.set_attr_string("fnetout", "mode", mode)
        o = "fnetout/output"
        f.create_group(o)
        f.set_attr_int(o, "ndatapoints", ndp)
        f.set_attr_int(o, "tforces", 1 if forces is not None else 0)
        f.set_attr_int(o, "nglobaltargets", n_global)
        f.set_attr_int(o, "natomictargets", n_atomic)
        for s in range(ndp):
            d = "%s/datapoint%d" % (o, s + 1)
            f.create_group(d)
            a, b = int(off[s]), int(off[s + 1])
            if n_global > 0:
                f.write_double(d + "/globalpredictions", _sum_atoms(raw[a:b, :n_global]))
                if mode == "validate":
                    f.write_double(d + "/globaltargets", np.asarray(global_targets)[s])
            f.write_double(d + "/rawpredictions", raw[a:b])
            if mode == "validate" and n_atomic > 0:
                f.write_double(d + "/atomictargets", np.asarray(atomic_targets)[a:b])
            if forces is not None:
                f.write_double(d + "/forces", np.asarray(forces, np.float64)[a:b])


def read_fnetout(path):
    """The arrays of an fnetout file: ``mode``, ``raw`` [nAtoms, nOut] over all datapoints,
    ``offsets``, ``glob`` [nStruct, nGlobal] when present, ``forces`` when ``tforces``."""
    with H5File(path) as f:
        mode = f.attr("fnetout", "mode").strip()
        o = "fnetout/output"
        ndp = int(f.attr(o, "ndatapoints"))
        tforces = int(f.attr(o, "tforces", 0))
        raw, glob, forces, n = [], [], [], [0]
        for i in range(1, ndp + 1):
            d = "%s/datapoint%d" % (o, i)
            r = f.dataset(d + "/rawpredictions", np.float64)
            r = r.reshape(r.shape[0], -1)
            raw.append(r)
            n.append(n[-1] + r.shape[0])
            if f.exists(d + "/globalpredictions"):
                glob.append(f.dataset(d + "/globalpredictions", np.float64).reshape(-1))
            if tforces:
                fr = f.dataset(d + "/forces", np.float64)
                forces.append(fr.reshape(fr.shape[0], -1))
    res = {"mode": mode, "raw": np.concatenate(raw), "offsets": np.array(n, np.int64),
           "tforces": tforces, "nglobaltargets": int(0 if not glob else len(glob[0]))}
    if glob:
        res["glob"] = np.stack(glob)
    if forces:
        res["forces"] = np.concatenate(forces)
    return res


# ---------------------------------------------------------------------------------------------
# a prediction run

def predict(dataset_path, netstat_path, fnetout_path, mode="predict", forces=True,
            precision=None, device=0, ctx=None):
    """Reads a dataset and a netstat, evaluates every structure on the device and writes the
    fnetout, like the reference's prediction and validation runs with analytic forces
    (``fortnet.F90`` predict / validate, ``forces.F90:317-425``, ``writeFnetout``).  Returns the
    arrays written (raw predictions, forces or None)."""
    from ._lib import FNET_FP64
    ds = read_dataset(dataset_path)
    ns = load_netstat(netstat_path)
    geos = _dataset.geometries(ds)
    own = ctx is None
    if own:
        ctx = Context(device)
    try:
        ctx.upload_model(ns)
        b = ctx.batch(geos, FNET_FP64 if precision is None else precision)
        try:
            _, raw = b.energy()
            F = b.forces() if forces else None
        finally:
            b.close()
    finally:
        if own:
            ctx.close()
    n_atomic = ns.n_out - ns.n_global
    write_fnetout(fnetout_path, mode, raw, ds["offsets"], ns.n_global, n_atomic, F,
                  ds["globaltargets"], ds["atomictargets"])
    return raw, F
