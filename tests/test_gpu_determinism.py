"""Reproducibility and robustness of the force and gradient assembly on the device.

* Forces are assembled without atomics (per-entry adjoint gradients gathered per atom through the
  mirror entries of the list, in row order) and weight gradients are reduced from per-tile
  partials in tile order, so two runs give bit-identical energies, forces and gradients.
* A pair exactly at the cutoff can be in one row and not in its mirror row: the reference's FP64
  test coords(j) + L n - coords(i) (dynneighlist.F90:291-294) rounds differently from the mirror
  expression.  Such an entry has no mirror; its neighbour side still reaches the forces, which
  equal the oracle's (which follows the reference's rows).
* New lattice vectors for a resident batch pass the same checks as a batch upload
  (fnetdata.F90:1151-1154; bounded lattice-image ranges) before anything is copied.
"""
import math

import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def _zscored(ctx, name, scale):
    geos, ns, _ = W.config(name, scale)
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    return geos, ns


@pytest.mark.parametrize("precision", [fa.FNET_FP64, fa.FNET_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("launch", [fa.FNET_BATCHED, fa.FNET_PER_STRUCT], ids=["batched", "per_struct"])
@pytest.mark.parametrize("name,scale", [("water", 0.004), ("si_cells", 0.004)])
def test_repeated_runs_are_bit_identical(ctx, name, scale, precision, launch):
    print("workload", name, scale, "precision", precision, "launch", launch, flush=True)
    geos, ns = _zscored(ctx, name, scale)
    outs = []
    for rep in range(2):
        b = ctx.batch(geos, precision, launch)
        b.run()
        Es, Ea = b.energy()
        F = b.forces()
        rng = np.random.default_rng(3)
        gt = rng.normal(size=(b.S, 1)) * 0.1 + Es[:, :1]
        grad, pred = b.train_gradients(gt)
        outs.append((Es, Ea, F, grad, pred, b.loss))
        b.close()
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(np.asarray(a), np.asarray(c))


def _asymmetric_pair():
    """Two atoms on the x axis of a 9-Bohr periodic direction and a cutoff between the two
    roundings of their image distance: row i holds (j, image +x), row j lacks (i, image -x)."""
    rng = np.random.default_rng(0)
    L = 9.0
    for _ in range(100000):
        xi, xj = rng.uniform(0.0, L, 2)
        d_row_i = (xi - (xj + L)) ** 2        # centre i, neighbour j shifted by +L
        d_row_j = (xj - (xi + (-L))) ** 2     # centre j, neighbour i shifted by -L
        if d_row_i < d_row_j:
            rc = math.sqrt(d_row_i)
            for _ in range(8):
                if rc * rc < d_row_i:
                    rc = float(np.nextafter(rc, np.inf))
                elif rc * rc >= d_row_j:
                    rc = float(np.nextafter(rc, -np.inf))
            if d_row_i <= rc * rc < d_row_j and abs(xi - xj) < rc:
                return xi, xj, L, rc
    raise AssertionError("no asymmetric pair found")


def test_forces_with_a_pair_at_the_cutoff(ctx):
    xi, xj, L, rc = _asymmetric_pair()
    print("asymmetric pair", xi, xj, L, rc, flush=True)
    rad = fa.GFunctions([2] * 3, [rc] * 3, eta=[0.01, 0.1, 0.5])
    ang = fa.GFunctions([5] * 4, [rc] * 4, eta=[0.02] * 4, xi=[1, 2, 4, 16],
                        lam=[1.0, -1.0, 1.0, -1.0])
    g = fa.GFunctions.concat([rad, ang])
    dims = [len(g), 5, 1]
    ns = fa.Netstat(g, [14], dims, 0, W.random_params(dims, 1, 7, scale=0.4),
                    zmean=np.zeros(len(g)), zstd=np.ones(len(g)))
    ctx.upload_model(ns)
    # a third atom off the axis gives the angular functions pairs
    xyz = np.array([[xi, 1.0, 1.0], [xj, 1.0, 1.0], [0.5 * (xi + xj), 3.0, 1.5]])
    lat = np.diag([L, 20.0, 20.0])
    geo = fa.Geometry(xyz, [14, 14, 14], [0, 0, 0], lat)
    b = ctx.batch([geo], fa.FNET_FP64)
    b.neighbour_list()
    rp, j, img = b.neighbours()
    rows = [set(zip(j[rp[a]:rp[a + 1]].tolist(), img[rp[a]:rp[a + 1]].tolist())) for a in range(3)]
    nimages = 3 * 3 * 3 - 1
    # row 0 holds atom 1 at an image whose mirror row 1 lacks
    missing = [(jj, k) for jj, k in rows[0] if jj == 1 and k >= 0 and (0, nimages - 1 - k) not in rows[1]]
    assert missing, "the test pair is symmetric in the device list"
    F = b.forces()
    Es, Ea = b.energy()
    gf = O.GFunctions(ns.g.type, ns.g.rcut, ns.g.eta, ns.g.rs, ns.g.kappa, ns.g.lam, ns.g.xi,
                      ns.g.atomid, ns.g.z1, ns.g.z2)
    st = O.Structure(xyz, lat, 1, np.array([14, 14, 14]))
    X = O.acsf(st, gf)
    nets = O.Nets(ns.dims, ns.activation, ns.params)
    J = nets.jacobian(X, np.zeros(3, np.int32))
    F0 = O.acsf_adjoint(st, gf, J[:, 0, :])
    E0 = nets.predict(X, np.zeros(3, np.int32))
    assert np.allclose(Ea, E0, atol=1e-10, rtol=1e-9)
    assert np.allclose(F, F0, atol=1e-10, rtol=1e-9), np.abs(F - F0).max()
    # the asymmetric entry makes the forces not sum to zero; both sides agree on it
    assert abs(F[:, 0].sum() - F0[:, 0].sum()) < 1e-10
    b.close()


def test_update_coords_rejects_bad_cells(ctx):
    geos, ns = _zscored(ctx, "si64", 0.02)
    b = ctx.batch(geos, fa.FNET_FP64)
    F_ref = b.forces()
    coords = np.concatenate([g.coords for g in geos])
    lats = np.stack([g.latvecs for g in geos])
    singular = lats.copy()
    singular[0, 2] = singular[0, 0] + singular[0, 1]      # dependent lattice vectors
    with pytest.raises(fa.FnetError, match="dependent lattice vectors"):
        b.update_coords(coords, singular)
    tiny = lats.copy()
    tiny[-1] *= 1e-3                                       # far more than 1e6 images
    with pytest.raises(fa.FnetError, match="cell too small"):
        b.update_coords(coords, tiny)
    bad = lats.copy()
    bad[0, 0, 0] = np.nan
    with pytest.raises(fa.FnetError, match="non-finite"):
        b.update_coords(coords, bad)
    # nothing was copied: the batch still evaluates its own structures
    F = b.forces()
    assert np.array_equal(F, F_ref)
    b.update_coords(coords, lats)
    assert np.array_equal(b.forces(), F_ref)
    b.close()


def test_results_follow_parameter_and_model_uploads(ctx):
    """A batch evaluated before new parameters (or a new model) are uploaded recomputes its
    outputs and forces when they are next read: they equal those of a fresh batch."""
    geos, ns = _zscored(ctx, "water", 0.004)
    b = ctx.batch(geos, fa.FNET_FP64)
    E0, _ = b.energy()
    b.forces()
    p2 = ns.params * 1.1
    ctx.upload_params(p2)
    E1, Ea1 = b.energy()
    F1 = b.forces()
    fresh = ctx.batch(geos, fa.FNET_FP64)
    E2, Ea2 = fresh.energy()
    F2 = fresh.forces()
    fresh.close()
    assert not np.allclose(E1, E0)
    assert np.array_equal(Ea1, Ea2) and np.array_equal(F1, F2)
    # a model with other functions (same counts) changes the features too
    g = ns.g
    g2 = fa.GFunctions(g.type, g.rcut * 0.9, eta=g.eta, rs=g.rs, kappa=g.kappa, lam=g.lam,
                       xi=g.xi, atomid=g.atomid, z1=g.z1, z2=g.z2)
    ns2 = fa.Netstat(g2, ns.species_z, ns.dims, ns.activation, p2, ns.zmean, ns.zstd)
    ctx.upload_model(ns2)
    G3 = b.acsf()
    E3, _ = b.energy()
    fresh = ctx.batch(geos, fa.FNET_FP64)
    assert np.array_equal(G3, fresh.acsf())
    assert np.array_equal(E3, fresh.energy()[0])
    fresh.close()
    b.close()


def test_element_slots_follow_model_uploads(ctx):
    """A batch keeps the atomic numbers of its atoms: after a model with another element table is
    uploaded, its atoms' element slots are derived again, and its features, energies and forces
    equal those of a fresh batch.  Si cells under a Si-only model (one element slot that every
    atom takes: the staging without the element sort), then under a Si/O model whose radial
    functions count Si and O neighbours apart (Si takes slot 1 of [O, Si], so slots kept from the
    first model would count the Si neighbours as O), then under the first model again."""
    geos, ns = _zscored(ctx, "si64", 0.03)
    b = ctx.batch(geos, fa.FNET_FP64)
    G0, E0, F0 = b.acsf(), b.energy()[0], b.forces()
    rc = float(ns.g.rcut[0])
    g = fa.GFunctions([2, 2, 2], [rc] * 3, eta=[0.01, 0.05, 0.01], z1=[14, 14, 8])
    dims = [3, 4, 1]
    ns2 = fa.Netstat(g, [14, 8], dims, 0, W.random_params(dims, 2, 7))
    ctx.upload_model(ns2)
    for prec in (fa.FNET_FP64, fa.FNET_FP32):
        b.set_mode(prec)
        G1, E1, F1 = b.acsf(), b.energy()[0], b.forces()
        fresh = ctx.batch(geos, prec)
        assert np.array_equal(G1, fresh.acsf())
        assert np.array_equal(E1, fresh.energy()[0])
        assert np.array_equal(F1, fresh.forces())
        fresh.close()
        assert (G1[:, 0] > 0).all() and (G1[:, 2] == 0).all()  # Si neighbours only
    ctx.upload_model(ns)
    b.set_mode(fa.FNET_FP64)
    assert np.array_equal(G0, b.acsf())
    assert np.array_equal(E0, b.energy()[0]) and np.array_equal(F0, b.forces())
    b.close()
