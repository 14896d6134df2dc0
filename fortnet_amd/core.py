"""Host-side mirror of Fortnet's hot-path interface over the HIP C-ABI.

The classes keep the reference's data model and semantics:

* :class:`GFunctions` -- the ACSF table of ``TGFunctions`` (``acsf.F90:79-91``), with the
  automatic parameter scheme ``TGFunctions_fromAutoScheme`` (``acsf.F90:276-363``) and the
  species-resolved expansion of ``initprogram.F90:1489-1524``;
* :class:`Netstat` -- the contents of a netstat file: G-functions, z-score statistics
  (``acsf.F90:445-536``), per-element networks (``TBpnn``, ``bpnn.F90:46-90``) with parameters
  serialised as ``network.F90:397-427``;
* :class:`Geometry` -- one ``TGeometry`` (``typegeometry.F90:24-59``): Cartesian coordinates in
  Bohr, lattice vectors as rows, atomic numbers and network (species) indices;
* :class:`Context` / :class:`Batch` -- the device state (``fnet_ctx`` / ``fnet_batch``).

Every computation runs in libfnet_hip.so; nothing here computes descriptors or networks.
"""
import atexit
import copy
import ctypes
import math
import weakref

import numpy as np

from . import _lib
from ._lib import FNET_BATCHED, FNET_FP32, FNET_FP64, FNET_PER_STRUCT, check, ptr

BOHR__AA = 0.529177249          # lib_dftbp/constants.F90:18
AA__BOHR = 1.0 / BOHR__AA       # lib_dftbp/constants.F90:21

G_TYPES = {"g1": 1, "g2": 2, "g3": 3, "g4": 4, "g5": 5}
ACTIVATIONS = ["tanh", "sigmoid", "relu", "lrelu", "softplus", "bent", "atan", "gaussian",
               "heaviside", "linear"]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class GFunctions:
    """Table of G-functions in netstat order (one row per function)."""

    def __init__(self, type, rcut, eta=None, rs=None, kappa=None, lam=None, xi=None, atomid=None,
                 z1=None, z2=None):
        n = len(type)
        zero = np.zeros(n)
        izero = np.zeros(n, np.int32)
        self.type = _i32(type)
        self.rcut = _f64(rcut)
        self.eta = _f64(zero if eta is None else eta)
        self.rs = _f64(zero if rs is None else rs)
        self.kappa = _f64(zero if kappa is None else kappa)
        self.lam = _f64(zero if lam is None else lam)
        self.xi = _f64(zero if xi is None else xi)
        self.atomid = _i32(izero if atomid is None else atomid)
        self.z1 = _i32(izero if z1 is None else z1)
        self.z2 = _i32(izero if z2 is None else z2)

    def __len__(self):
        return len(self.type)

    @classmethod
    def concat(cls, tables):
        keys = ["type", "rcut", "eta", "rs", "kappa", "lam", "xi", "atomid", "z1", "z2"]
        return cls(*[np.concatenate([getattr(t, k) for t in tables]) for k in keys])

    @classmethod
    def auto(cls, rcut, n_radial, n_angular, atomid=0):
        """``TGFunctions_fromAutoScheme`` (acsf.F90:276-363); rcut in Bohr."""
        rs_step = rcut / (n_radial - 1)
        g2eta = 5.0 * math.log(10.0) / (2.0 * rs_step) ** 2
        g2rs = [0.0] + [ii * rs_step for ii in range(1, n_radial)]
        g5eta = 2.0 * math.log(10.0) / rcut ** 2
        lam, xis = [], []
        for ii in range(0, math.ceil(n_angular / 2.0 - 1.0) + 1):
            xi = 1.0 if n_angular <= 2 else 1.0 + ii * 30.0 / (n_angular - 2.0)
            for jj in (1, -1):
                if len(lam) < n_angular:
                    lam.append(float(jj))
                    xis.append(xi)
        types = [2] * n_radial + [5] * n_angular
        return cls(types, [rcut] * (n_radial + n_angular),
                   eta=[g2eta] * n_radial + [g5eta] * n_angular,
                   rs=g2rs[:n_radial] + [0.0] * n_angular,
                   lam=[0.0] * n_radial + lam, xi=[0.0] * n_radial + xis,
                   atomid=[atomid] * (n_radial + n_angular))

    def species_resolved(self, atomic_numbers):
        """Expansion of ``initprogram.F90:1489-1524``: radial functions once per species
        (atomicNumbers = (Z, 0)), angular functions once per combination with replacement
        (Z_a, Z_b), a <= b, in dataset order."""
        rows = []
        zs = list(atomic_numbers)
        pairs = [(zs[a], zs[b]) for a in range(len(zs)) for b in range(a, len(zs))]
        for f in range(len(self)):
            base = {k: getattr(self, k)[f] for k in
                    ("type", "rcut", "eta", "rs", "kappa", "lam", "xi", "atomid")}
            if self.type[f] <= 3:
                for z in zs:
                    rows.append(dict(base, z1=z, z2=0))
            else:
                for z1, z2 in pairs:
                    rows.append(dict(base, z1=z1, z2=z2))
        keys = ["type", "rcut", "eta", "rs", "kappa", "lam", "xi", "atomid", "z1", "z2"]
        return GFunctions(*[[r[k] for r in rows] for k in keys])


class Netstat:
    """Model state: ACSF table, z-score, external inputs and per-element networks."""

    def __init__(self, gfuncs, species_z, dims, activation, params, zmean=None, zstd=None,
                 ext_in=None, n_global=None):
        self.g = gfuncs
        self.species_z = _i32(species_z)
        self.dims = _i32(dims)
        nsp = len(self.species_z)
        self.activation = _i32(np.broadcast_to(np.asarray(activation), (nsp,)))
        self.params = _f64(params).reshape(nsp, -1)
        self.zmean = None if zmean is None else _f64(zmean)
        self.zstd = None if zstd is None else _f64(zstd)
        self.ext_in = _i32([] if ext_in is None else ext_in)
        self.n_global = int(self.dims[-1] if n_global is None else n_global)
        if self.params.shape[1] != nparams(self.dims):
            raise ValueError("parameter vector length %d != %d" % (self.params.shape[1],
                                                                   nparams(self.dims)))

    @property
    def n_out(self):
        return int(self.dims[-1])

    @classmethod
    def from_arrays(cls, d, prefix=""):
        """From the arrays of a converted netstat file (keys as written by the fixture export)."""
        g = lambda k: d[prefix + "g_" + k]
        gf = None
        if prefix + "g_type" in d:
            gf = GFunctions(g("type"), g("rcut"), g("eta"), g("rs"), g("kappa"), g("lam"),
                            g("xi"), g("atomid"), g("z1"), g("z2"))
        else:
            gf = GFunctions([], [])
        return cls(gf, d[prefix + "net_atomicnumbers"], d[prefix + "topology"],
                   d[prefix + "activation"], d[prefix + "params"],
                   d.get(prefix + "zmean"), d.get(prefix + "zstd"), d.get(prefix + "ext_indices"),
                   int(d[prefix + "nglobaltargets"]))


def nparams(dims):
    """nWeights + nBiases of ``TNetwork_countParams`` (network.F90:126-142)."""
    dims = [int(x) for x in dims]
    nw = sum(dims[l] * dims[l + 1] for l in range(len(dims) - 1)) + dims[-1]
    return nw + sum(dims)


class Geometry:
    """One structure: coords [n,3] in Bohr (unfolded), latvecs rows = lattice vectors."""

    def __init__(self, coords, atnum, species, latvecs=None, ext=None):
        self.coords = _f64(coords).reshape(-1, 3)
        self.atnum = _i32(atnum)
        self.species = _i32(species)
        self.periodic = latvecs is not None
        self.latvecs = _f64(latvecs if latvecs is not None else np.zeros((3, 3))).reshape(3, 3)
        n = len(self.coords)
        self.ext = _f64(ext if ext is not None else np.zeros((n, 0))).reshape(n, -1)

    @property
    def n(self):
        return self.coords.shape[0]


_live_contexts = weakref.WeakSet()
_shutdown = [False]


@atexit.register
def _close_all():
    """Releases device state while the HIP runtime is still alive; later finalisers do nothing."""
    for c in list(_live_contexts):
        c.close()
    _shutdown[0] = True


class Context:
    """A device context (``fnet_ctx``): one HIP stream and the resident model."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.fnet_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise _lib.FnetError("fnet_ctx_create(%d) failed with status %d (no usable GPU?)"
                                 % (device, rc))
        self.h = h
        self.model = None
        self._batches = weakref.WeakSet()
        _live_contexts.add(self)

    def close(self):
        """Destroys the context and every batch created on it."""
        if _shutdown[0]:
            return
        for b in list(self._batches):
            b.close()
        if self.h:
            self.lib.fnet_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_model(self, ns):
        """``fnet_model_upload``: ACSF table, z-score and networks stay resident."""
        g = ns.g
        self._keep = [g, ns]
        d = _lib.ModelDesc()
        d.n_g = len(g)
        d.g_type, d.g_rcut, d.g_eta = ptr(g.type), ptr(g.rcut), ptr(g.eta)
        d.g_rs, d.g_kappa, d.g_lambda, d.g_xi = ptr(g.rs), ptr(g.kappa), ptr(g.lam), ptr(g.xi)
        d.g_atomid, d.g_z1, d.g_z2 = ptr(g.atomid), ptr(g.z1), ptr(g.z2)
        d.z_mean = ptr(ns.zmean)
        d.z_std = ptr(ns.zstd)
        d.n_ext_in = len(ns.ext_in)
        d.ext_in_idx = ptr(ns.ext_in) if len(ns.ext_in) else None
        d.n_species = len(ns.species_z)
        d.species_z = ptr(ns.species_z)
        d.n_layers = len(ns.dims)
        d.dims = ptr(ns.dims)
        d.activation = ptr(ns.activation)
        d.params = ptr(ns.params)
        check(self.h, self.lib.fnet_model_upload(self.h, ctypes.byref(d)), "fnet_model_upload")
        # the context's own record of the resident model: parameter updates (optimizer steps,
        # upload_params) change it, never the caller's netstat
        self.model = copy.copy(ns)
        self.model.params = np.array(ns.params, np.float64)

    def upload_params(self, params):
        params = _f64(params).reshape(self.model.params.shape)
        check(self.h, self.lib.fnet_params_upload(self.h, ptr(params)), "fnet_params_upload")
        self.model.params = params.copy()

    def batch(self, geos, precision=FNET_FP32, launch=FNET_BATCHED):
        return Batch(self, geos, precision, launch)

    def set_loss(self, loss="mse"):
        """Loss of the training gradient: 'mse', 'mae', 'mape' or 'rms' (``fnet_set_loss``)."""
        check(self.h, self.lib.fnet_set_loss(self.h, _lib.LOSSES[loss]), "fnet_set_loss")

    def set_regularization(self, kind=None, strength=0.0, alpha=0.0):
        """Elastic-net regularisation of the weight gradients (``fnet_set_regularization``):
        kind 'lasso' (alpha 1), 'ridge' (alpha 0), 'elasticnet' (alpha given) or None."""
        if kind is None or strength == 0.0:
            strength, alpha = 0.0, 0.0
        elif kind == "lasso":
            alpha = 1.0
        elif kind == "ridge":
            alpha = 0.0
        elif kind != "elasticnet":
            raise ValueError("invalid regularisation type %r" % kind)
        check(self.h, self.lib.fnet_set_regularization(self.h, float(strength), float(alpha)),
              "fnet_set_regularization")

    def set_graph(self, enable=True):
        """``fnet_set_graph``: fnet_run replays a captured HIP graph of its steady-state step."""
        check(self.h, self.lib.fnet_set_graph(self.h, 1 if enable else 0), "fnet_set_graph")

    def device_memory(self):
        """(free, total) device memory in bytes (``fnet_device_memory``)."""
        f, t = ctypes.c_longlong(0), ctypes.c_longlong(0)
        check(self.h, self.lib.fnet_device_memory(self.h, ctypes.byref(f), ctypes.byref(t)),
              "fnet_device_memory")
        return int(f.value), int(t.value)

    def step_time(self):
        """Device time (ms) of the last fnet_run / fnet_forces step (``fnet_step_time``)."""
        ms = ctypes.c_double(0.0)
        check(self.h, self.lib.fnet_step_time(self.h, ctypes.byref(ms)), "fnet_step_time")
        return ms.value

    def optimizer_init(self, kind, tol=1e-20, min_disp=1e-6, max_disp=1e4, learning_rate=0.01,
                       memory=1000, line_search=True):
        """``fnet_optimizer_init``: device FP64 optimizer ('sd', 'cg', 'lbfgs', 'fire') per
        element network, started at the resident parameters."""
        self._opt = _lib.OPTIMIZERS[kind.lower()]
        check(self.h, self.lib.fnet_optimizer_init(self.h, self._opt, float(tol), float(min_disp),
                                                   float(max_disp), float(learning_rate),
                                                   int(memory), 1 if line_search else 0),
              "fnet_optimizer_init")

    def optimizer_step(self):
        """``fnet_optimizer_step`` on the gradient of the last train call: (new params,
        converged flag of the last network)."""
        out = np.zeros_like(self.model.params)
        conv = ctypes.c_int(0)
        check(self.h, self.lib.fnet_optimizer_step(self.h, self._opt, ptr(out),
                                                   ctypes.byref(conv)), "fnet_optimizer_step")
        self.model.params = out
        return out, bool(conv.value)

    def sd_step(self, lr, max_disp=1e4, tol=0.0):
        """``fnet_optimizer_step_sd`` on the gradient of the last train call."""
        out = np.zeros_like(self.model.params)
        conv = ctypes.c_int(0)
        check(self.h, self.lib.fnet_optimizer_step_sd(self.h, lr, max_disp, tol, ptr(out),
                                                      ctypes.byref(conv)), "fnet_optimizer_step_sd")
        self.model.params = out
        return out, bool(conv.value)


class Batch:
    """A resident batch of structures (``fnet_batch``)."""

    def __init__(self, ctx, geos, precision=FNET_FP32, launch=FNET_BATCHED):
        self.ctx = ctx
        lib = ctx.lib
        self.lib = lib
        d = self._describe(geos)
        h = ctypes.c_void_p()
        check(ctx.h, lib.fnet_batch_upload(ctx.h, ctypes.byref(d), ctypes.byref(h)),
              "fnet_batch_upload")
        self.h = h
        ctx._batches.add(self)
        self.set_mode(precision, launch)

    def _describe(self, geos):
        """Packs the geometries and returns the ``fnet_batch_desc`` pointing at them."""
        self.S = len(geos)
        n = [g.n for g in geos]
        self.offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(n)]), np.int64)
        self.N = int(self.offsets[-1])
        self.coords = _f64(np.concatenate([g.coords for g in geos]) if geos else np.zeros((0, 3)))
        self.latvecs = _f64(np.stack([g.latvecs for g in geos]) if geos else np.zeros((0, 3, 3)))
        self.periodic = _i32([int(g.periodic) for g in geos])
        self.species = _i32(np.concatenate([g.species for g in geos]) if geos else [])
        self.atnum = _i32(np.concatenate([g.atnum for g in geos]) if geos else [])
        next_ = max([g.ext.shape[1] for g in geos] + [0])
        self.n_ext = next_
        self.ext = None
        if next_ > 0:
            self.ext = _f64(np.concatenate([g.ext for g in geos]))
        d = _lib.BatchDesc()
        d.n_struct = self.S
        d.offsets = ptr(self.offsets)
        d.coords = ptr(self.coords)
        d.latvecs = ptr(self.latvecs)
        d.periodic = ptr(self.periodic)
        d.species = ptr(self.species)
        d.atnum = ptr(self.atnum)
        d.n_ext = next_
        d.ext = ptr(self.ext)
        return d

    def reload(self, geos):
        """``fnet_batch_reload``: new structures in the same device buffers (serving)."""
        d = self._describe(geos)
        check(self.ctx.h, self.lib.fnet_batch_reload(self.h, ctypes.byref(d)), "fnet_batch_reload")

    def query(self):
        """True while work queued on the context is still running (``fnet_query``)."""
        rc = self.lib.fnet_query(self.ctx.h)
        if rc not in (0, 1):
            check(self.ctx.h, rc, "fnet_query")
        return rc == 1

    def close(self):
        if _shutdown[0]:
            return
        if self.h and self.ctx.h:
            self.lib.fnet_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, precision, launch=FNET_BATCHED):
        check(self.ctx.h, self.lib.fnet_batch_set_mode(self.h, precision, launch),
              "fnet_batch_set_mode")
        self.precision = precision
        self.launch = launch

    def update_coords(self, coords, latvecs=None):
        coords = _f64(coords).reshape(self.N, 3)
        lat = None if latvecs is None else _f64(latvecs).reshape(self.S, 3, 3)
        check(self.ctx.h, self.lib.fnet_batch_update_coords(self.h, ptr(coords), ptr(lat)),
              "fnet_batch_update_coords")

    def neighbour_list(self):
        n = ctypes.c_longlong(0)
        check(self.ctx.h, self.lib.fnet_neighbour_list(self.ctx.h, self.h, ctypes.byref(n)),
              "fnet_neighbour_list")
        self.npairs = int(n.value)
        return self.npairs

    def neighbours(self):
        """(rowptr [N+1], j [P], img [P]) of the last neighbour list."""
        rp = np.zeros(self.N + 1, np.int64)
        j = np.zeros(max(self.npairs, 1), np.int32)
        img = np.zeros(max(self.npairs, 1), np.int32)
        check(self.ctx.h, self.lib.fnet_get_neighbours(self.ctx.h, self.h, ptr(rp), ptr(j),
                                                       ptr(img)), "fnet_get_neighbours")
        return rp, j[:self.npairs], img[:self.npairs]

    def acsf(self):
        """Unstandardised ACSF values [N, nG] (``TAcsf%calculate``)."""
        G = np.zeros((self.N, len(self.ctx.model.g)))
        check(self.ctx.h, self.lib.fnet_acsf_forward(self.ctx.h, self.h, ptr(G)),
              "fnet_acsf_forward")
        return G

    def energy(self):
        """(E_struct [S, nOut], E_atom [N, nOut])."""
        no = self.ctx.model.n_out
        Es = np.zeros((self.S, no))
        Ea = np.zeros((self.N, no))
        check(self.ctx.h, self.lib.fnet_energy(self.ctx.h, self.h, ptr(Es), ptr(Ea)), "fnet_energy")
        return Es, Ea

    def forces(self):
        """F [N, 3*nOut] in the fnetout layout (F[k, 3t + c])."""
        F = np.zeros((self.N, 3 * self.ctx.model.n_out))
        check(self.ctx.h, self.lib.fnet_forces(self.ctx.h, self.h, ptr(F)), "fnet_forces")
        return F

    def run(self):
        check(self.ctx.h, self.lib.fnet_run(self.ctx.h, self.h), "fnet_run")

    TUNE_KNOBS = ("fwd_wpb", "adj_wpb", "mlp_waves", "rbal", "persist")

    def set_tune(self, fwd_wpb=0, adj_wpb=0, mlp_waves=0, rbal=-1, persist=0):
        """Launch parameters of this batch's model, precision and size (``fnet_set_tune``)."""
        k = np.array([fwd_wpb, adj_wpb, mlp_waves, rbal, persist], np.int32)
        check(self.ctx.h, self.lib.fnet_set_tune(self.ctx.h, self.h, ptr(k)), "fnet_set_tune")

    def get_tune(self):
        k = np.zeros(5, np.int32)
        check(self.ctx.h, self.lib.fnet_get_tune(self.ctx.h, self.h, ptr(k)), "fnet_get_tune")
        return dict(zip(self.TUNE_KNOBS, (int(x) for x in k)))

    def autotune(self, reps=3):
        """``fnet_autotune``: chooses the launch parameters by timing the batch's steady-state
        step; returns (parameters, step ms with the defaults, step ms tuned)."""
        ms = np.zeros(2)
        k = np.zeros(5, np.int32)
        check(self.ctx.h, self.lib.fnet_autotune(self.ctx.h, self.h, int(reps), ptr(ms), ptr(k)),
              "fnet_autotune")
        return dict(zip(self.TUNE_KNOBS, (int(x) for x in k))), float(ms[0]), float(ms[1])

    def sync(self):
        """Waits for the batch's last step and validates it (``fnet_batch_sync``)."""
        check(self.ctx.h, self.lib.fnet_batch_sync(self.h), "fnet_batch_sync")

    def graph_replayed(self):
        """True if the last ``run`` replayed the batch's HIP graph."""
        g = ctypes.c_int(0)
        check(self.ctx.h, self.lib.fnet_batch_step_info(self.h, ctypes.byref(g)),
              "fnet_batch_step_info")
        return bool(g.value)

    def stage_times(self):
        ms = np.zeros(8)
        check(self.ctx.h, self.lib.fnet_stage_times(self.ctx.h, ptr(ms), 5), "fnet_stage_times")
        return ms[:5]

    def train_gradients(self, g_targets=None, a_targets=None, atomic_weights=None,
                        dp_weights=None, n_global=None):
        m = self.ctx.model
        ng = m.n_global if n_global is None else n_global
        gt = None if g_targets is None or ng == 0 else _f64(g_targets).reshape(self.S, ng)
        at = None
        if m.n_out - ng > 0:
            at = _f64(a_targets).reshape(self.N, m.n_out - ng)
        aw = None if atomic_weights is None else _f64(atomic_weights).reshape(self.N)
        dw = None if dp_weights is None else _f64(dp_weights).reshape(self.S)
        grad = np.zeros_like(m.params)
        pred = np.zeros((self.N, m.n_out))
        loss = np.zeros(1)
        check(self.ctx.h, self.lib.fnet_train_gradients(self.ctx.h, self.h, ng, ptr(gt), ptr(at),
                                                        ptr(aw), ptr(dw), ptr(grad), ptr(pred),
                                                        ptr(loss)),
              "fnet_train_gradients")
        self.loss = float(loss[0])
        return grad, pred
