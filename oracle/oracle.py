"""Python front-end of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker.  The product
(``fortnet_amd``) never imports it.

It wraps ``fnet_oracle.c`` (a plain-C restatement of the reference's hot path)
through ctypes and adds the host-side glue of the reference's driver:

* fractional -> Cartesian coordinates, ``fnetdata.F90:1147-1150``;
* z-score of values and derivatives, ``acsf.F90:489-536`` (skip if sigma < 1e-8);
* network input = ACSF values (+ external features), ``features.F90:200-265``;
* predictions / global sums / analytic forces, ``forces.F90:317-425``;
* central finite-difference forces, ``forces.F90:182-313``.

Every function is pinned by the reference fixtures, closed forms, invariants or finite
differences (tests/test_oracle.py); the ten activations by the scratch-training fixtures
(tests/test_oracle_training.py).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libfnet_oracle.so")
SRC = os.path.join(HERE, "fnet_oracle.c")

REFSHAPE = 1
COMPAT_SCATTER = 2
COMPAT_SELFIMG = 4
UNREG_COS = 8  # planted mistake for the tests: cos(theta) without the 1e-13 regulariser

ACTIVATIONS = ["tanh", "sigmoid", "relu", "lrelu", "softplus", "bent", "atan", "gaussian",
               "heaviside", "linear"]

BOHR__AA = 0.529177249  # lib_dftbp/constants.F90:18


def build(force=False):
    """Compile the oracle with gcc (no FMA contraction, no fast-math)."""
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp%d" % os.getpid()
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                           "-std=c99", "-o", tmp, SRC, "-lm"])
    os.replace(tmp, LIB)
    return LIB


_lib = None


class _GF(ctypes.Structure):
    _fields_ = [("nG", ctypes.c_int), ("type", ctypes.c_void_p), ("rcut", ctypes.c_void_p),
                ("eta", ctypes.c_void_p), ("rs", ctypes.c_void_p), ("kappa", ctypes.c_void_p),
                ("lam", ctypes.c_void_p), ("xi", ctypes.c_void_p), ("atomid", ctypes.c_void_p),
                ("z1", ctypes.c_void_p), ("z2", ctypes.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(build())
        _lib.orc_nparams.restype = ctypes.c_long
        _lib.orc_neighbours.restype = ctypes.c_long
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class GFunctions:
    """G-function table in the netstat mapping layout (``acsf.F90:40-75``)."""

    def __init__(self, type, rcut, eta=None, rs=None, kappa=None, lam=None, xi=None, atomid=None,
                 z1=None, z2=None):
        n = len(type)
        z = np.zeros(n)
        self.type = np.ascontiguousarray(type, np.int32)
        self.rcut = np.ascontiguousarray(rcut, np.float64)
        self.eta = np.ascontiguousarray(z if eta is None else eta, np.float64)
        self.rs = np.ascontiguousarray(z if rs is None else rs, np.float64)
        self.kappa = np.ascontiguousarray(z if kappa is None else kappa, np.float64)
        self.lam = np.ascontiguousarray(z if lam is None else lam, np.float64)
        self.xi = np.ascontiguousarray(z if xi is None else xi, np.float64)
        zi = np.zeros(n, np.int32)
        self.atomid = np.ascontiguousarray(zi if atomid is None else atomid, np.int32)
        self.z1 = np.ascontiguousarray(zi if z1 is None else z1, np.int32)
        self.z2 = np.ascontiguousarray(zi if z2 is None else z2, np.int32)

    @property
    def n(self):
        return len(self.type)

    def c(self):
        return _GF(self.n, _p(self.type), _p(self.rcut), _p(self.eta), _p(self.rs), _p(self.kappa),
                   _p(self.lam), _p(self.xi), _p(self.atomid), _p(self.z1), _p(self.z2))

    @classmethod
    def from_case(cls, case, prefix="in_"):
        g = lambda k: case[prefix + "g_" + k]
        return cls(g("type"), g("rcut"), g("eta"), g("rs"), g("kappa"), g("lam"), g("xi"),
                   g("atomid"), g("z1"), g("z2"))


class Structure:
    """One geometry: Cartesian coordinates in Bohr, lattice rows = lattice vectors."""

    def __init__(self, coords, lat, periodic, atnum, ext=None):
        self.coords = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
        self.lat = np.ascontiguousarray(lat if lat is not None else np.zeros((3, 3)), np.float64)
        self.periodic = int(periodic)
        self.atnum = np.ascontiguousarray(atnum, np.int32)
        n = self.coords.shape[0]
        self.ext = np.ascontiguousarray(ext if ext is not None else np.zeros((n, 0)), np.float64)

    @property
    def n(self):
        return self.coords.shape[0]


def frac_to_cart(frac, lat):
    """``matmul(latVecs, frac)`` (``fnetdata.F90:1147-1148``), accumulated in the same order."""
    frac = np.asarray(frac, np.float64)
    out = np.zeros_like(frac)
    for r in range(3):
        acc = np.zeros(frac.shape[0])
        for c in range(3):
            acc = acc + lat[c, r] * frac[:, c]
        out[:, r] = acc
    return out


def dataset_structures(ds):
    """Structures of a converted fnetdata file."""
    out = []
    off = ds["offsets"]
    for s in range(int(ds["ndatapoints"])):
        a, b = int(off[s]), int(off[s + 1])
        xyz = ds["coords"][a:b]
        lat = ds["basis"][s]
        if ds["periodic"][s] and ds["fractional"][s]:
            xyz = frac_to_cart(xyz, lat)
        out.append(Structure(xyz, lat, ds["periodic"][s], ds["atnum"][a:b], ds["ext"][a:b]))
    return out


def neighbours(st, rc, i):
    """Reference-order neighbour entries of centre ``i``: (j, image index, image coords, dist^2)."""
    L = lib()
    cap = 4096
    while True:
        j = np.zeros(cap, np.int32)
        img = np.zeros(cap, np.int32)
        xyz = np.zeros((cap, 3))
        d2 = np.zeros(cap)
        c = L.orc_neighbours(ctypes.c_int(st.n), _p(st.coords), _p(st.lat), ctypes.c_int(st.periodic),
                             ctypes.c_double(rc), ctypes.c_int(i), ctypes.c_long(cap), _p(j), _p(img),
                             _p(xyz), _p(d2))
        if c <= cap:
            return j[:c], img[:c], xyz[:c], d2[:c]
        cap = c


def image_ranges(lat, rc):
    lo = np.zeros(3, np.int32)
    hi = np.zeros(3, np.int32)
    lib().orc_image_ranges(_p(np.ascontiguousarray(lat, np.float64)), ctypes.c_double(rc), _p(lo),
                           _p(hi))
    return lo, hi


def acsf(st, gf, flags=0):
    G = np.zeros((st.n, gf.n))
    gfc = gf.c()
    lib().orc_acsf(ctypes.c_int(st.n), _p(st.coords), _p(st.lat), ctypes.c_int(st.periodic),
                   _p(st.atnum), _p(st.ext), ctypes.c_int(st.ext.shape[1]), ctypes.byref(gfc),
                   ctypes.c_int(flags), _p(G))
    return G


def acsf_prime(st, gf, flags=0):
    """Dense derivatives dG[k, i, a, c] = d G_{i,a} / d r_{k,c}."""
    dG = np.zeros((st.n, st.n, gf.n, 3))
    gfc = gf.c()
    lib().orc_acsf_prime(ctypes.c_int(st.n), _p(st.coords), _p(st.lat), ctypes.c_int(st.periodic),
                         _p(st.atnum), _p(st.ext), ctypes.c_int(st.ext.shape[1]),
                         ctypes.byref(gfc), ctypes.c_int(flags), _p(dG))
    return dG


def acsf_adjoint(st, gf, cG, flags=0):
    """F[k] = -sum_{i,a} cG[i,a] dG_{i,a}/dr_k without the dense array."""
    cG = np.ascontiguousarray(cG, np.float64)
    F = np.zeros((st.n, 3))
    gfc = gf.c()
    lib().orc_acsf_adjoint(ctypes.c_int(st.n), _p(st.coords), _p(st.lat), ctypes.c_int(st.periodic),
                           _p(st.atnum), _p(st.ext), ctypes.c_int(st.ext.shape[1]),
                           ctypes.byref(gfc), ctypes.c_int(flags), _p(cG), _p(F))
    return F


class Nets:
    """Per-species networks in the serialised netstat layout (``network.F90:397-427``)."""

    def __init__(self, dims, act, params):
        self.dims = np.ascontiguousarray(dims, np.int32)
        self.act = np.ascontiguousarray(np.broadcast_to(act, (params.shape[0],)), np.int32)
        self.params = np.ascontiguousarray(params, np.float64)
        assert self.params.shape[1] == lib().orc_nparams(len(self.dims), _p(self.dims))

    @classmethod
    def from_case(cls, case, prefix="in_"):
        return cls(case[prefix + "topology"], case[prefix + "activation"], case[prefix + "params"])

    @property
    def nparams(self):
        return self.params.shape[1]

    def predict(self, feats, species):
        feats = np.ascontiguousarray(feats, np.float64)
        species = np.ascontiguousarray(species, np.int32)
        out = np.zeros((feats.shape[0], int(self.dims[-1])))
        lib().orc_bpnn_predict(ctypes.c_int(feats.shape[0]), ctypes.c_int(len(self.dims)),
                               _p(self.dims), _p(self.act), _p(self.params),
                               ctypes.c_long(self.nparams), _p(species), _p(feats), _p(out))
        return out

    def jacobian(self, feats, species):
        feats = np.ascontiguousarray(feats, np.float64)
        species = np.ascontiguousarray(species, np.int32)
        J = np.zeros((feats.shape[0], int(self.dims[-1]), int(self.dims[0])))
        lib().orc_bpnn_jacobian(ctypes.c_int(feats.shape[0]), ctypes.c_int(len(self.dims)),
                                _p(self.dims), _p(self.act), _p(self.params),
                                ctypes.c_long(self.nparams), _p(species), _p(feats), _p(J))
        return J

    def sys_train(self, feats, species, nglobal, gtargets, atargets, atomicweights, dpweight, grad):
        """Accumulate one structure's gradient into ``grad`` (bpnn.F90:436-451,610-704)."""
        feats = np.ascontiguousarray(feats, np.float64)
        species = np.ascontiguousarray(species, np.int32)
        nout = int(self.dims[-1])
        pred = np.zeros((feats.shape[0], nout))
        gt = np.ascontiguousarray(gtargets if gtargets is not None else np.zeros(1), np.float64)
        at = np.ascontiguousarray(atargets if atargets is not None else np.zeros(1), np.float64)
        aw = np.ascontiguousarray(atomicweights, np.float64)
        assert grad.flags["C_CONTIGUOUS"] and grad.dtype == np.float64
        lib().orc_sys_train(ctypes.c_int(feats.shape[0]), ctypes.c_int(len(self.dims)),
                            _p(self.dims), _p(self.act), _p(self.params),
                            ctypes.c_long(self.nparams), _p(species), _p(feats),
                            ctypes.c_int(nglobal), _p(gt), _p(at), _p(aw),
                            ctypes.c_double(dpweight), _p(pred), _p(grad))
        return pred


def zscore(G, zmean, zstd):
    """``TAcsf_applyZscore`` (acsf.F90:489-508)."""
    G = G.copy()
    if zmean is None:
        return G
    for a in range(G.shape[1]):
        if zstd[a] < 1e-8:
            continue
        G[:, a] = (G[:, a] - zmean[a]) / zstd[a]
    return G


def zscale(zstd, nG):
    """Per-function factor of ``TAcsf_applyZscorePrime`` (acsf.F90:512-536)."""
    s = np.ones(nG)
    if zstd is None:
        return s
    for a in range(nG):
        if zstd[a] >= 1e-8:
            s[a] = 1.0 / zstd[a]
    return s


def case_features(case, st, gf, flags=0):
    """Network input of one structure: z-scored ACSF (+ external features by index)."""
    parts = []
    if gf is not None:
        G = acsf(st, gf, flags)
        zm = case["in_zmean"] if "in_zmean" in case else None
        zs = case["in_zstd"] if "in_zstd" in case else None
        parts.append(zscore(G, zm, zs))
    if "in_ext_indices" in case:
        idx = case["in_ext_indices"]
        parts.append(st.ext[:, idx - 1])
    return np.concatenate(parts, axis=1)


def predict_case(case, ds, flags=0, forces=True):
    """Predictions and analytic forces for every structure of a converted test case.

    Returns (raw[Natoms, nOut], glob[S, nGlobal], forces[Natoms, 3*nOut] or None), the layout of
    the reference's fnetout file.
    """
    gf = GFunctions.from_case(case) if "in_g_type" in case else None
    nets = Nets.from_case(case)
    nglob = int(case["in_nglobaltargets"])
    structs = dataset_structures(ds)
    off = ds["offsets"]
    raws, globs, frcs = [], [], []
    for s, st in enumerate(structs):
        species = ds["globalsp"][off[s]:off[s + 1]] - 1
        X = case_features(case, st, gf, flags)
        raw = nets.predict(X, species)
        raws.append(raw)
        globs.append(raw[:, :nglob].sum(axis=0))
        if forces and gf is not None:
            J = nets.jacobian(X, species)  # [n, nOut, nIn]
            sc = zscale(case["in_zstd"] if "in_zstd" in case else None, gf.n)
            nout = J.shape[1]
            F = np.zeros((st.n, 3 * nout))
            for t in range(nout):
                cG = J[:, t, :gf.n] * sc[None, :]
                F[:, t * 3:(t + 1) * 3] = acsf_adjoint(st, gf, cG, flags)
            frcs.append(F)
    return (np.concatenate(raws), np.stack(globs),
            np.concatenate(frcs) if frcs else None)


def dense_forces(case, st, species, flags=0):
    """Analytic forces through the dense dG array (``forces.F90:400-413``)."""
    gf = GFunctions.from_case(case)
    nets = Nets.from_case(case)
    X = case_features(case, st, gf, flags)
    J = nets.jacobian(X, species)
    dG = acsf_prime(st, gf, flags) * zscale(case.get("in_zstd"), gf.n)[None, None, :, None]
    nout = J.shape[1]
    F = np.zeros((st.n, 3 * nout))
    for t in range(nout):
        # F[k, c] = - sum_i sum_a J[i,t,a] dG[k,i,a,c]
        F[:, 3 * t:3 * t + 3] = -np.einsum("ia,kiac->kc", J[:, t, :gf.n], dG)
    return F


def energy(case, st, species, flags=0):
    gf = GFunctions.from_case(case)
    nets = Nets.from_case(case)
    X = case_features(case, st, gf, flags)
    return nets.predict(X, species).sum(axis=0)


def fd_forces(case, st, species, delta=0.01, flags=0):
    """Central finite differences of the summed outputs (``forces.F90:182-313``)."""
    nout = int(case["in_topology"][-1])
    F = np.zeros((st.n, 3 * nout))
    for k in range(st.n):
        for c in range(3):
            xp = st.coords.copy()
            xm = st.coords.copy()
            xp[k, c] += delta
            xm[k, c] -= delta
            ep = energy(case, Structure(xp, st.lat, st.periodic, st.atnum, st.ext), species, flags)
            em = energy(case, Structure(xm, st.lat, st.periodic, st.atnum, st.ext), species, flags)
            F[k, c::3] = (em - ep) / (2.0 * delta)
    return F
