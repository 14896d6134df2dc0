"""Edge structures on the device path against the oracle, in one batch and one per launch.

* a lone atom and a pair of atoms beyond each other's cutoff: rows with no neighbour (all
  features zero, forces exactly zero);
* a one-atom periodic cell smaller than the cutoff: the atom's only neighbours are its own
  images (the j = k and self-image terms of acsf.F90:1377-1492), and its force is zero;
* a two-atom periodic cell and a small cluster, next to them: ragged rows in one batch;
* species-resolved lists where an element is absent around a centre: empty element ranges of
  radial functions and of both kinds of angular groups;
* rows longer than 64 KiB of LDS staging take a block of their own (gfx950 gives one block up to
  160 KiB): a three-atom cell of 1.7 A^3 whose centres have ~900 neighbour images each;
* a structure too dense even for that: the host reports it as an error before any ACSF kernel
  runs (every kernel that did run, the neighbour search, ends normally).

FP64 is held to 1e-10 relative to the largest value (as the other oracle comparisons), FP32 to
the accuracy gates (1e-5 eV/atom, force RMSE 1e-4 eV/Angstrom).
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HARTREE__EV = 27.211386245988
BOHR__AA = fa.BOHR__AA


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def _oracle_all(geos, ns):
    gf = O.GFunctions(ns.g.type, ns.g.rcut, ns.g.eta, ns.g.rs, ns.g.kappa, ns.g.lam, ns.g.xi,
                      ns.g.atomid, ns.g.z1, ns.g.z2)
    nets = O.Nets(ns.dims, ns.activation, ns.params)
    Gs, Ea, Fs = [], [], []
    for g in geos:
        st = O.Structure(g.coords, g.latvecs, int(g.periodic), g.atnum, g.ext)
        G = O.acsf(st, gf)
        X = O.zscore(G, ns.zmean, ns.zstd)
        Ea.append(nets.predict(X, g.species))
        J = nets.jacobian(X, g.species)
        sc = O.zscale(ns.zstd, gf.n)
        Fs.append(O.acsf_adjoint(st, gf, J[:, 0, :] * sc[None, :]))
        Gs.append(G)
    return np.concatenate(Gs), np.concatenate(Ea), np.concatenate(Fs)


def _model():
    rc = 5.0 * fa.AA__BOHR
    rad = fa.GFunctions([2] * 3, [rc] * 3, eta=W.eta_bohr([0.01, 0.1, 1.0]))
    g4 = fa.GFunctions([4] * 4, [rc] * 4, eta=W.eta_bohr([0.005] * 4), xi=[1, 1, 4, 4],
                       lam=[1.0, -1.0, 1.0, -1.0])
    g5 = fa.GFunctions([5] * 4, [rc] * 4, eta=W.eta_bohr([0.02] * 4), xi=[1, 2, 16, 8],
                       lam=[1.0, -1.0, 1.0, -1.0])
    g = fa.GFunctions.concat([rad, g4, g5]).species_resolved([1, 8, 14])
    dims = [len(g), 6, 1]
    ns = fa.Netstat(g, [1, 8, 14], dims, 0, W.random_params(dims, 3, 21, scale=0.3),
                    zmean=np.zeros(len(g)), zstd=np.ones(len(g)))
    return ns, rc


def _geos(rc):
    A = fa.AA__BOHR
    sp = {1: 0, 8: 1, 14: 2}

    def geo(xyz, z, lat=None):
        z = np.asarray(z)
        return fa.Geometry(np.asarray(xyz, float) * A, z, [sp[int(x)] for x in z],
                           None if lat is None else np.asarray(lat, float) * A)

    rng = np.random.default_rng(5)
    cl = rng.uniform(-1.6, 1.6, size=(7, 3))
    return [
        geo([[0.0, 0.0, 0.0]], [8]),                                   # lone atom
        geo([[0.0, 0.0, 0.0], [0.0, 0.0, 6.0]], [1, 14]),              # beyond each other's rc
        geo([[0.3, 0.2, 0.1]], [14], np.eye(3) * 2.6),                # self-images only
        geo([[0.0, 0.0, 0.0], [1.2, 1.1, 0.9]], [14, 1],
            [[3.1, 0.0, 0.0], [0.4, 3.3, 0.0], [0.0, 0.2, 3.5]]),      # two-atom triclinic cell
        geo(cl, [1, 8, 1, 1, 8, 1, 1]),                               # cluster, no Si around
    ]


@pytest.mark.parametrize("launch", [fa.FNET_BATCHED, fa.FNET_PER_STRUCT],
                         ids=["batched", "per_struct"])
def test_edge_structures_match_oracle(ctx, launch):
    ns, rc = _model()
    geos = _geos(rc)
    ctx.upload_model(ns)
    G0, E0, F0 = _oracle_all(geos, ns)
    # the rows without neighbours really have none
    assert np.all(G0[0] == 0) and np.all(G0[1] == 0) and np.all(G0[2] == 0)
    for prec in (fa.FNET_FP64, fa.FNET_FP32):
        print("precision", prec, "launch", launch, flush=True)
        b = ctx.batch(geos, prec)
        b.set_mode(prec, launch)
        G = b.acsf()
        _, Ea = b.energy()
        F = b.forces()
        b.close()
        assert np.all(F[:3] == 0.0)  # no neighbours: no force at all
        assert np.abs(F[3]).max() < 1e-9 * max(1.0, np.abs(F0).max()) or prec == fa.FNET_FP32
        if prec == fa.FNET_FP64:
            assert np.abs(G - G0).max() <= 1e-10 * max(1.0, np.abs(G0).max())
            assert np.abs(Ea - E0).max() <= 1e-10 * max(1.0, np.abs(E0).max())
            assert np.abs(F - F0).max() <= 1e-10 * max(1.0, np.abs(F0).max())
        else:
            assert np.abs(Ea - E0).max() * HARTREE__EV <= 1e-5
            rmse = np.sqrt(np.mean((F - F0) ** 2)) * HARTREE__EV / BOHR__AA
            assert rmse <= 1e-4


def test_long_rows_take_a_block_of_their_own(ctx):
    """Three Si atoms in a triclinic cell of 1.7 A^3 and a 5 A cutoff: rows of 899-909 neighbour
    images, whose per-centre LDS staging (FP32 adjoint with paired G5 groups: 73 KiB; FP64:
    117 KiB) exceeds 64 KiB and runs one wave per block; next to it a two-atom cell with short
    rows.  The G5 functions form two groups that differ only in eta (the paired walk)."""
    A = fa.AA__BOHR
    rc = 5.0 * A
    rad = fa.GFunctions([2] * 2, [rc] * 2, eta=W.eta_bohr([0.05, 0.3]))
    g5 = fa.GFunctions([5] * 4, [rc] * 4, eta=W.eta_bohr([0.02, 0.02, 0.06, 0.06]),
                       xi=[1, 4, 1, 4], lam=[1.0, -1.0, 1.0, -1.0])
    g = fa.GFunctions.concat([rad, g5])
    dense = fa.Geometry(np.array([[0.1, 0.2, 0.15], [0.6, 0.5, 0.4], [0.3, 0.9, 0.8]]) * A,
                        [14] * 3, [0] * 3,
                        np.array([[1.15, 0.0, 0.0], [0.1, 1.2, 0.0], [0.05, 0.1, 1.25]]) * A)
    pair = fa.Geometry(np.array([[0.0, 0.0, 0.0], [1.3, 1.1, 0.9]]) * A, [14] * 2, [0] * 2,
                       np.array([[3.4, 0.0, 0.0], [0.3, 3.6, 0.0], [0.0, 0.2, 3.8]]) * A)
    geos = [dense, pair]
    dims = [len(g), 8, 1]
    ns = fa.Netstat(g, [14], dims, 0, W.random_params(dims, 1, 17, scale=0.4),
                    zmean=np.zeros(len(g)), zstd=np.ones(len(g)))
    G0, _, _ = _oracle_all(geos, ns)
    ns = fa.Netstat(g, [14], dims, 0, ns.params, zmean=np.zeros(len(g)),
                    zstd=np.abs(G0).max(axis=0))          # features of order one
    _, E0, F0 = _oracle_all(geos, ns)
    ctx.upload_model(ns)
    for prec, launch in ((fa.FNET_FP64, fa.FNET_BATCHED), (fa.FNET_FP32, fa.FNET_BATCHED),
                         (fa.FNET_FP64, fa.FNET_PER_STRUCT)):
        print("precision", prec, "launch", launch, flush=True)
        b = ctx.batch(geos, prec, launch)
        b.neighbour_list()
        rp, _, _ = b.neighbours()
        assert np.diff(rp)[:3].min() >= 899
        G = b.acsf()
        _, Ea = b.energy()
        F = b.forces()
        b.close()
        if prec == fa.FNET_FP64:
            assert np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()
            assert np.abs(Ea - E0).max() <= 1e-10 * max(1.0, np.abs(E0).max())
            assert np.abs(F - F0).max() <= 1e-10 * max(1.0, np.abs(F0).max())
        else:
            # each dense feature sums ~400 000 non-negative FP32 terms (~6 000 per lane), so
            # the features are held to 1e-4 relative and the energies to 1e-4 of their
            # scale; the forces to the force gate
            grel = (np.abs(G - G0) / np.abs(G0)).max()
            erel = np.abs(Ea - E0).max() / np.abs(E0).max()
            rmse = np.sqrt(np.mean((F - F0) ** 2)) * HARTREE__EV / BOHR__AA
            print("FP32 feature error", grel, "energy error", erel, "force RMSE eV/A", rmse,
                  flush=True)
            assert grel <= 1e-4 and erel <= 1e-4 and rmse <= 1e-4


def test_too_dense_structure_is_refused(ctx):
    """2 700 atoms within the cutoff of each other: rows of 2 699 entries exceed the adjoint's LDS
    staging of one centre (173 KiB in FP32 against the 160 KiB one block may take), which the
    host checks before the ACSF kernels run."""
    ns, rc = _model()
    ctx.upload_model(ns)
    rng = np.random.default_rng(3)
    n = 2700
    xyz = rng.uniform(0.0, 2.5, size=(n, 3)) * fa.AA__BOHR  # box diagonal < rc
    geo = fa.Geometry(xyz, np.full(n, 1), np.zeros(n, np.int32))
    b = ctx.batch([geo], fa.FNET_FP32)
    with pytest.raises(fa.FnetError, match="too many neighbours around one atom"):
        b.energy()
        b.forces()
    b.close()
