"""GPU parity on the synthetic workloads (small instances) against the FP64 oracle.

* FP64 path: ACSF values, energies and forces within 1e-10 (relative to the largest value);
* FP32 path: the accuracy gates of the batched mixed-precision path -- energy error
  <= 1e-5 eV/atom and force RMSE <= 1e-4 eV/Angstrom (network outputs are read as eV, lengths
  are Bohr, so forces are converted with 1 eV/Bohr = 1/0.529177249 eV/Angstrom);
* per-structure launches and the batched launch give the same numbers;
* training gradients: FP64 within 1e-9 relative, FP32 within 1e-3 relative (stated, not gated
  by the reference).
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

E_TOL = 1e-5            # eV/atom
F_RMSE_TOL = 1e-4       # eV/Angstrom
EV_BOHR_TO_EV_AA = 1.0 / fa.BOHR__AA


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def _oracle_all(geos, ns):
    """Oracle G, atomic outputs, forces for a list of geometries."""
    gf = O.GFunctions(ns.g.type, ns.g.rcut, ns.g.eta, ns.g.rs, ns.g.kappa, ns.g.lam, ns.g.xi,
                      ns.g.atomid, ns.g.z1, ns.g.z2)
    nets = O.Nets(ns.dims, ns.activation, ns.params)
    Gs, Ea, Fs = [], [], []
    for g in geos:
        st = O.Structure(g.coords, g.latvecs, int(g.periodic), g.atnum, g.ext)
        G = O.acsf(st, gf)
        X = O.zscore(G, ns.zmean, ns.zstd)
        out = nets.predict(X, g.species)
        J = nets.jacobian(X, g.species)
        sc = O.zscale(ns.zstd, gf.n)
        F = O.acsf_adjoint(st, gf, J[:, 0, :] * sc[None, :])
        Gs.append(G)
        Ea.append(out)
        Fs.append(F)
    return np.concatenate(Gs), np.concatenate(Ea), np.concatenate(Fs)


CASES = {
    "si64": ("si64", 0.02),
    "water": ("water", 0.004),
    "si_cells": ("si_cells", 0.0015),
    "si_bulk": ("si_bulk", 0.125),
    "ternary": ("ternary", 0.016),
}


def _setup(ctx, name):
    cfg, scale = CASES[name] if name in CASES else CASES_FP64[name]
    geos, ns, _ = W.config(cfg, scale)
    if name == "ternary":
        geos = [g for g in geos if g.n <= 400][:2] or geos[:1]
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    return geos, ns


CASES_FP64 = dict(CASES, si_bulk_cells=("si_bulk", 0.25))  # 2048 atoms: cell-list search


@pytest.mark.parametrize("name", list(CASES_FP64))
def test_fp64_matches_oracle(ctx, name):
    print("workload", name, flush=True)
    geos, ns = _setup(ctx, name)
    G0, E0, F0 = _oracle_all(geos, ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    G = b.acsf()
    Es, Ea = b.energy()
    F = b.forces()
    assert np.abs(G - G0).max() <= 1e-10 * max(1.0, np.abs(G0).max())
    assert np.abs(Ea - E0).max() <= 1e-10 * max(1.0, np.abs(E0).max())
    assert np.abs(F - F0).max() <= 1e-9 * max(1.0, np.abs(F0).max())
    off = np.concatenate([[0], np.cumsum([g.n for g in geos])])
    Es0 = np.array([E0[off[s]:off[s + 1]].sum() for s in range(len(geos))])
    assert np.abs(Es[:, 0] - Es0).max() <= 1e-10 * max(1.0, np.abs(Es0).max())
    # per-structure launches (the plain path) give the same numbers
    b.set_mode(fa.FNET_FP64, fa.FNET_PER_STRUCT)
    Es2, Ea2 = b.energy()
    F2 = b.forces()
    assert np.abs(Ea2 - Ea).max() <= 1e-12 * max(1.0, np.abs(Ea).max())
    assert np.abs(F2 - F).max() <= 1e-11 * max(1.0, np.abs(F).max())
    b.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [fa.FNET_FP32, fa.FNET_FP32_NET64], ids=["fp32", "fp32net64"])
def test_fp32_meets_accuracy_gates(ctx, name, precision):
    """Absolute gates on every workload.  The perturbed lattices (si64, si_bulk) standardise
    radial features whose spread between atoms is 0.8-2 % of their mean: the z-score divides
    G - mean by that sigma, so the FP32 radial sums' own error (~4e-7 relative) reached the network
    input 50-120 times larger and moved the si_bulk forces by 1.6e-4 eV/A RMS (measured with a
    per-term FP32 restatement on the CPU; the device gave 1.9e-4).  Such models (|mean| / sigma
    above 25 for some radial feature) have their radial sums evaluated in FP64; the FP32 angular
    sums, the FP32 adjoint and the FP32 pair vectors then leave ~6e-5 eV/A."""
    print("workload", name, "precision", precision, flush=True)
    geos, ns = _setup(ctx, name)
    b = ctx.batch(geos, fa.FNET_FP64)
    Es64, _ = b.energy()
    F64 = b.forces()
    b.set_mode(precision)
    Es32, _ = b.energy()
    F32 = b.forces()
    nat = np.array([g.n for g in geos])
    e_err = np.abs(Es32[:, 0] - Es64[:, 0]) / nat
    f_rmse = np.sqrt(np.mean((F32 - F64) ** 2)) * EV_BOHR_TO_EV_AA
    print("%s precision %d: max |dE|/atom %.3e eV, force RMSE %.3e eV/A"
          % (name, precision, e_err.max(), f_rmse))
    b.close()
    assert e_err.max() <= E_TOL
    assert f_rmse <= F_RMSE_TOL


@pytest.mark.parametrize("name", ["si64", "water"])
def test_train_gradients_match_oracle(ctx, name):
    print("workload", name, flush=True)
    geos, ns = _setup(ctx, name)
    gf = O.GFunctions(ns.g.type, ns.g.rcut, ns.g.eta, ns.g.rs, ns.g.kappa, ns.g.lam, ns.g.xi,
                      ns.g.atomid, ns.g.z1, ns.g.z2)
    nets = O.Nets(ns.dims, ns.activation, ns.params)
    rng = np.random.default_rng(0)
    targets = rng.normal(size=(len(geos), 1)) * 3.0
    weights = rng.integers(1, 4, size=len(geos)).astype(np.float64)
    aw = np.concatenate([rng.uniform(0.5, 1.5, size=g.n) for g in geos])
    grad0 = np.zeros_like(ns.params)
    off = 0
    for s, g in enumerate(geos):
        st = O.Structure(g.coords, g.latvecs, int(g.periodic), g.atnum)
        X = O.zscore(O.acsf(st, gf), ns.zmean, ns.zstd)
        nets.sys_train(X, g.species, 1, targets[s], None, aw[off:off + g.n], weights[s], grad0)
        off += g.n
    grad0 /= weights.sum()
    for prec, tol in ((fa.FNET_FP64, 1e-9), (fa.FNET_FP32, 1e-3)):
        b = ctx.batch(geos, prec)
        grad, _ = b.train_gradients(targets, None, aw, weights)
        rel = np.abs(grad - grad0).max() / np.abs(grad0).max()
        print("%s precision %d: gradient max rel err %.3e" % (name, prec, rel))
        assert rel <= tol
        b.close()


@pytest.mark.parametrize("name,scale", [("si_cells", 0.004), ("si_bulk", 0.125),
                                        ("ternary", 0.03), ("water", 0.004)])
def test_train_gradients_precisions(ctx, name, scale):
    """Weight gradients of the FP32 and BF16 modes (matrix-core kernel: 64-atom blocks on the
    24-64-64-1 and 48-64-64-1 networks, 32-atom halves of a tile on 32-128-128-1) against the
    FP64 gradients of the same batch: FP32 and split BF16 within 1e-3 of the largest component,
    single-pass BF16 within 5e-2 (reported, not gated).  Repeated calls give the same bits."""
    print("workload", name, scale, flush=True)
    geos, ns, _ = W.config(name, scale)
    if name == "ternary":
        geos = [g for g in geos if g.n <= 600][:4]
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    rng = np.random.default_rng(1)
    b = ctx.batch(geos, fa.FNET_FP64)
    Es, _ = b.energy()
    targets = Es[:, :1] + rng.normal(size=(len(geos), 1)) * np.array([[g.n for g in geos]]).T * 1e-2
    g64, _ = b.train_gradients(targets)
    scale_g = np.abs(g64).max()
    errs = {}
    for prec, tol in ((fa.FNET_FP32, 1e-3), (fa.FNET_BF16, 1e-3), (fa.FNET_BF16_FAST, 5e-2)):
        b.set_mode(prec)
        g1, _ = b.train_gradients(targets)
        g2, _ = b.train_gradients(targets)
        assert np.array_equal(g1, g2)
        errs[prec] = np.abs(g1 - g64).max() / scale_g
        assert errs[prec] <= tol, (prec, errs[prec])
    print("%s gradient max err / max |g|: fp32 %.2e bf16 %.2e bf16_fast %.2e"
          % (name, errs[fa.FNET_FP32], errs[fa.FNET_BF16], errs[fa.FNET_BF16_FAST]))
    b.close()


@pytest.mark.parametrize("name", ["si64", "water", "ternary"])
def test_generic_exponent_path(ctx, name):
    """One more angular function with xi = 3 (same type, cutoff, eta, species and atom id as the
    last one, so it joins that group) takes the model off the exponent-ladder kernels (every xi a
    power of two) onto the per-member loop.  The shared features must agree between the two
    kernels, and the generic model must match the oracle in FP64 and meet the FP32 gates."""
    print("workload", name, flush=True)
    geos, ns = _setup(ctx, name)
    b = ctx.batch(geos, fa.FNET_FP64)
    G_lad = b.acsf()
    b.set_mode(fa.FNET_FP32)
    G_lad32 = b.acsf()
    b.close()
    g = ns.g
    k = int(np.nonzero((g.type == 4) | (g.type == 5))[0][-1])
    extra = fa.GFunctions([g.type[k]], [g.rcut[k]], eta=[g.eta[k]], lam=[g.lam[k]], xi=[3.0],
                          atomid=[g.atomid[k]], z1=[g.z1[k]], z2=[g.z2[k]])
    g2 = fa.GFunctions.concat([g, extra])
    dims = ns.dims.copy()
    dims[0] += 1
    ns2 = fa.Netstat(g2, ns.species_z, dims, ns.activation,
                     W.random_params(dims, len(ns.species_z), 7))
    ctx.upload_model(ns2)
    b = ctx.batch(geos, fa.FNET_FP64)
    G2 = b.acsf()
    b.set_mode(fa.FNET_FP32)
    G2_32 = b.acsf()
    b.close()
    scale = max(1.0, np.abs(G_lad).max())
    assert np.abs(G2[:, :-1] - G_lad).max() <= 1e-12 * scale
    assert np.abs(G2_32[:, :-1] - G_lad32).max() <= 1e-5 * scale
    ns2 = W.with_zscore(ns2, G2)
    ctx.upload_model(ns2)
    G0, E0, F0 = _oracle_all(geos, ns2)
    b = ctx.batch(geos, fa.FNET_FP64)
    G = b.acsf()
    _, Ea = b.energy()
    F = b.forces()
    assert np.abs(G - G0).max() <= 1e-10 * max(1.0, np.abs(G0).max())
    assert np.abs(Ea - E0).max() <= 1e-10 * max(1.0, np.abs(E0).max())
    assert np.abs(F - F0).max() <= 1e-9 * max(1.0, np.abs(F0).max())
    b.set_mode(fa.FNET_FP32)
    Es32, _ = b.energy()
    F32 = b.forces()
    nat = np.array([geo.n for geo in geos])
    Es64 = np.array([E0[o:o + n].sum() for o, n in zip(np.cumsum(nat) - nat, nat)])
    assert (np.abs(Es32[:, 0] - Es64) / nat).max() <= E_TOL
    assert np.sqrt(np.mean((F32 - F0) ** 2)) * EV_BOHR_TO_EV_AA <= F_RMSE_TOL
    b.close()


def _clusters(seed, sizes=range(2, 11)):
    """Non-periodic H/O clusters small enough that every atom sees every other one."""
    rng = np.random.default_rng(seed)
    geos = []
    for n in sizes:
        while True:
            xyz = rng.uniform(-2.2, 2.2, size=(n, 3))
            d = np.linalg.norm(xyz[:, None] - xyz[None], axis=2) + 10 * np.eye(n)
            if d.min() > 0.9:
                break
        z = rng.choice([1, 8], size=n)
        geos.append(fa.Geometry(xyz * fa.AA__BOHR, z, (z == 8).astype(np.int32)))
    return geos


@pytest.mark.parametrize("generic", [False, True], ids=["ladder", "generic"])
def test_small_clusters_match_oracle(ctx, generic):
    """Neighbour lists of 0-9 entries per element: rows of one or two entries, odd and even row
    lengths, species lists of different lengths in either order, for G4 and G5 groups; the
    two-pairs-per-step walk of the exponent-ladder kernels (and, with one xi = 3 member, the
    per-member walk) must give the oracle's values, energies and forces in FP64."""
    rc = 6.5 * fa.AA__BOHR
    geos = _clusters(11)
    rad = fa.GFunctions([2] * 3, [rc] * 3, eta=W.eta_bohr([0.01, 0.1, 1.0]))
    g4 = fa.GFunctions([4] * 4, [rc] * 4, eta=W.eta_bohr([0.005] * 4), xi=[1, 1, 4, 4],
                       lam=[1.0, -1.0, 1.0, -1.0])
    xi5 = [1, 2, 16, 3 if generic else 8]
    g5 = fa.GFunctions([5] * 4, [rc] * 4, eta=W.eta_bohr([0.02] * 4), xi=xi5,
                       lam=[1.0, -1.0, 1.0, -1.0])
    g = fa.GFunctions.concat([rad, g4, g5]).species_resolved([1, 8])
    dims = [len(g), 8, 1]
    ns = fa.Netstat(g, [1, 8], dims, 0, W.random_params(dims, 2, 9))
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    G0, E0, F0 = _oracle_all(geos, ns)
    for launch in (fa.FNET_BATCHED, fa.FNET_PER_STRUCT):
        b = ctx.batch(geos, fa.FNET_FP64)
        b.set_mode(fa.FNET_FP64, launch)
        G = b.acsf()
        _, Ea = b.energy()
        F = b.forces()
        b.close()
        assert np.abs(G - G0).max() <= 1e-10 * max(1.0, np.abs(G0).max())
        assert np.abs(Ea - E0).max() <= 1e-10 * max(1.0, np.abs(E0).max())
        assert np.abs(F - F0).max() <= 1e-9 * max(1.0, np.abs(F0).max())
