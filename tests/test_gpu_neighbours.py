"""Neighbour lists on the GPU are bit-exact with the reference's iterator.

The oracle restates ``TNeighIterator_getNextNeighbours`` (dynneighlist.F90:233-320) with the
reference's FP64 operation order; the GPU list must have the same entries -- neighbour index and
lattice image -- in the same order for every centre, including small cells where atoms meet
their own images and several images of one neighbour.
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import dataset
from fortnet_amd import workloads
from fortnet_amd import workloads as W
from conftest import load_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def _model(rc, species_z=(14,)):
    g = fa.GFunctions([2], [rc], eta=[0.1])
    dims = [1, 2, 1]
    nsp = len(species_z)
    return fa.Netstat(g, species_z, dims, 0, np.zeros((nsp, fa.nparams(dims))))


def _check_rows(geos, rc, rp, j, img):
    off = 0
    for g in geos:
        st = O.Structure(g.coords, g.latvecs, g.periodic, g.atnum)
        for i in range(g.n):
            oj, oimg, _, _ = O.neighbours(st, rc, i)
            a = off + i
            gj = j[rp[a]:rp[a + 1]] - off
            gi = img[rp[a]:rp[a + 1]]
            assert np.array_equal(gj, oj), (i, gj[:10], oj[:10])
            assert np.array_equal(gi, oimg), i
        off += g.n


def _compare(ctx, geos, rc):
    ctx.upload_model(_model(rc, sorted(set(int(z) for g in geos for z in g.atnum))))
    for prec in (fa.FNET_FP64, fa.FNET_FP32):
        b = ctx.batch(geos, prec)
        P = b.neighbour_list()
        rp, j, img = b.neighbours()
        _check_rows(geos, rc, rp, j, img)
        assert P == rp[-1]
        b.close()


def test_fixture_cells_exact(ctx):
    c, ds = load_case("analysis__derivatives__analytical__singleTarget__speciesResolved__raw")
    geos = dataset.geometries(ds)
    _compare(ctx, geos, float(c["in_g_rcut"].max()))


def test_cluster_exact(ctx):
    c, ds = load_case("validation__multispecies__bothTargets")
    _compare(ctx, dataset.geometries(ds)[:4], 7.5589039)


@pytest.mark.parametrize("seed", [0, 1])
def test_small_cells_with_self_images(ctx, seed):
    """1..70-atom cells shorter than the cutoff (many images per neighbour, self images)."""
    rng = np.random.default_rng(seed)
    geos = []
    for n in (1, 2, 8, 13, 40, 70):
        geos.append(workloads.random_cell(rng, n, 2.33, 2.0, 14))
    _compare(ctx, geos, 6.0 * fa.AA__BOHR)


def test_cell_list_path_exact(ctx):
    """Structures of >= 1024 atoms with 5 x 5 x 3 bins of width rc take the cell-list kernels; the
    sorted rows must still be the reference's entries in the reference's order, including atoms
    moved out of the cell by lattice vectors (unwrapped coordinates)."""
    g = W.si_diamond(6, 6, 4, 0.1, 11)
    assert g.n >= 1024
    xyz = g.coords.copy()
    L = g.latvecs
    xyz[::7] += L[0]
    xyz[3::11] -= L[2]
    g2 = fa.Geometry(xyz, g.atnum, g.species, L)
    _compare(ctx, [g, g2], 6.0 * fa.AA__BOHR)


def test_cell_list_mid_size_cells(ctx):
    """Cells of a few hundred atoms whose width allows exactly 3 bins per axis: the 27 stencil
    bins would cover the whole cell, which the direct scan does with fewer candidates, so these
    take the direct scan; cells with 3 bins on some axes only (below and the 1152-atom diamond
    above, 5 x 5 x 3 bins) take the cell lists, whose 3-bin axes get every image from the wrap."""
    rng = np.random.default_rng(3)
    geos = [workloads.random_cell(rng, n, 2.33, 2.0, 14) for n in (300, 420, 512)]
    geos.append(W.si_diamond(7, 3, 4, 0.1, 12))  # 6 x 2 x 3 bins: direct scan (2 on one axis)
    geos.append(W.si_diamond(8, 8, 4, 0.1, 13))  # 7 x 7 x 3 bins: cell lists
    _compare(ctx, geos, 6.0 * fa.AA__BOHR)


def test_mid_size_and_slab_cells(ctx):
    """Cells 12-18 A wide and slab / needle cells of the same volumes: fewer than 3 bins of
    width rc fit on some axis, so they take the direct scan over the reference's image ranges
    (a cell-list search for them, with the stencil meeting one bin under several wraps, measured
    1 ms slower on the 8192 Si cells); the rows must be the reference's entries in its order."""
    rng = np.random.default_rng(7)
    geos = [workloads.random_cell(rng, n, 2.33, 2.0, 14) for n in (90, 130, 200, 280)]
    for shape, n in (((2.1, 1.0, 1.0), 160), ((1.0, 2.6, 1.3), 200), ((3.2, 1.05, 1.1), 260)):
        g = workloads.random_cell(rng, n, 2.33, 2.0, 14)
        f = np.array(shape) / np.prod(shape) ** (1.0 / 3.0)  # same volume, other aspect
        geos.append(fa.Geometry(g.coords * f[None, :], g.atnum, g.species, g.latvecs * f[:, None]))
    _compare(ctx, geos, 6.0 * fa.AA__BOHR)


@pytest.mark.parametrize("precision", [fa.FNET_FP64, fa.FNET_FP32], ids=["fp64", "fp32"])
def test_one_pass_search_and_overflow(ctx, precision):
    """After the first list of a batch, batched lists come from one search pass into rows sized by
    the previous list, compacted afterwards.  Compressing the cells by 15 % gives rows longer than
    that: the batch must fall back to the counted fill, and every list must be the reference's."""
    rng = np.random.default_rng(5)
    geos = [workloads.random_cell(rng, n, 2.33, 2.0, 14) for n in (13, 40, 300)]
    geos.append(W.si_diamond(6, 6, 6, 0.1, 2))  # 1728 atoms: cell lists (5 and, compressed, 4 bins)
    rc = 6.0 * fa.AA__BOHR
    ctx.upload_model(_model(rc))
    b = ctx.batch(geos, precision)
    b.neighbour_list()
    first = b.neighbours()
    b.neighbour_list()
    second = b.neighbours()
    for x, y in zip(first, second):
        assert np.array_equal(x, y)
    _check_rows(geos, rc, *second)
    dense = [fa.Geometry(g.coords * 0.85, g.atnum, g.species, g.latvecs * 0.85) for g in geos]
    b.update_coords(np.concatenate([g.coords for g in dense]), np.stack([g.latvecs for g in dense]))
    for _ in range(2):  # overflow, then one pass with the enlarged rows
        P = b.neighbour_list()
        rp, j, img = b.neighbours()
        assert P == rp[-1]
        _check_rows(dense, rc, rp, j, img)
    b.close()


def test_triclinic_and_unwrapped(ctx):
    rng = np.random.default_rng(7)
    L = np.array([[9.0, 0.0, 0.0], [2.5, 8.0, 0.0], [-1.5, 3.0, 10.0]])
    xyz = rng.uniform(-0.3, 1.3, size=(30, 3)) @ L  # some atoms outside the cell
    g = fa.Geometry(xyz, np.full(30, 14), np.zeros(30, np.int32), L)
    _compare(ctx, [g], 7.0)


@pytest.mark.parametrize("precision", [fa.FNET_FP64, fa.FNET_FP32], ids=["fp64", "fp32"])
def test_cell_list_rows_beyond_64k_of_lds(ctx, precision):
    """A 20.5 x 15.5 x 15.5 A cell of 8 185 atoms and a 5 A cutoff (4 x 3 x 3 bins: the cell
    lists) with 767-962 neighbours per centre.  The first list sorts each centre's hits in LDS at a
    capacity of 1 024; the one-pass search of the next evaluation sorts at the row capacity it
    keeps (1.25 x the longest row), 2 048 entries: 73 KiB (FP32) or 106 KiB (FP64) of LDS, beyond
    64 KiB, in a block of its own.  Both lists are the reference's rows and give bit-identical
    energies and forces."""
    A = fa.AA__BOHR
    rc = 5.0 * A
    rng = np.random.default_rng(11)
    lat = np.diag([20.5, 15.5, 15.5]) * A
    n = 8185
    xyz = rng.uniform(0.0, 1.0, size=(n, 3)) @ lat
    geo = fa.Geometry(xyz, [14] * n, [0] * n, lat)
    rad = fa.GFunctions([2] * 2, [rc] * 2, eta=W.eta_bohr([0.05, 0.3]))
    g5 = fa.GFunctions([5] * 2, [rc] * 2, eta=W.eta_bohr([0.02, 0.06]), xi=[2, 2], lam=[1.0, 1.0])
    g = fa.GFunctions.concat([rad, g5])
    dims = [len(g), 4, 1]
    zstd = np.array([700.0, 70.0, 4e5, 6e4])  # features of order one
    ctx.upload_model(fa.Netstat(g, [14], dims, 0, W.random_params(dims, 1, 5, scale=0.4),
                                zmean=np.zeros(len(g)), zstd=zstd))
    b = ctx.batch([geo], precision)
    b.neighbour_list()                      # counted list: sort capacity 1 024
    rp1, j1, img1 = b.neighbours()
    rows = np.diff(rp1)
    print("precision", precision, "longest row", rows.max(), "mean", rows.mean(), flush=True)
    assert 820 <= rows.max() <= 1024
    E1, F1 = b.energy()[1], b.forces()
    b.neighbour_list()                      # one pass into the kept rows: capacity 2 048
    rp2, j2, img2 = b.neighbours()
    assert np.array_equal(rp1, rp2) and np.array_equal(j1, j2) and np.array_equal(img1, img2)
    b.update_coords(xyz, lat[None])         # the steady-state step
    E2, F2 = b.energy()[1], b.forces()
    assert np.array_equal(E1, E2) and np.array_equal(F1, F2)
    _check_rows([geo], rc, rp2, j2, img2)
    b.close()
