"""The C-ABI library (CPU-side checks, no kernel launches)."""
import ctypes
import os
import re

import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import _lib
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fnet_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fnet_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names


def test_hdf5_layer_exports_every_declared_function():
    """libfnet_h5.so (include/fnet_h5.h) and the ctypes binding of fortnet_amd/h5io.py."""
    from fortnet_amd import h5io
    if not os.path.exists(h5io.H5_LIB_PATH):
        pytest.skip("libfnet_h5.so not built (no HDF5 libraries)")
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "fnet_h5.h")).read(),
                 flags=re.S)
    names = sorted(set(re.findall(r"\b(fnet_h5_[a-z_0-9]+)\s*\(", txt)))
    assert len(names) == 21
    lib = h5io._load()
    for n in names:
        assert getattr(lib, n).argtypes is not None, n


def test_params_per_species_matches_reference_count():
    lib = _lib.load()
    for dims in ([16, 20, 20, 1], [32, 25, 25, 1], [18, 3, 3, 1], [22, 2, 2, 1], [5, 1]):
        d = np.asarray(dims, np.int32)
        assert lib.fnet_params_per_species(len(d), _lib.ptr(d)) == fa.nparams(dims)
    # network.F90:126-142: nWeights = sum d_l d_{l+1} + d_n, nBiases = sum d_l
    assert fa.nparams([16, 20, 20, 1]) == (16 * 20 + 20 * 20 + 20 * 1 + 1) + (16 + 20 + 20 + 1)


def test_context_without_gpu_fails_loudly():
    # the kernel driver's device node decides whether the HIP runtime can see a GPU
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(fa.FnetError):
        fa.Context(0)


def test_auto_scheme_and_species_expansion():
    """TGFunctions_fromAutoScheme (acsf.F90:276-363) and initprogram.F90:1489-1524 ordering."""
    rc = 4.0 * fa.AA__BOHR
    g = fa.GFunctions.auto(rc, 5, 4)
    step = rc / 4
    assert list(g.type) == [2] * 5 + [5] * 4
    assert np.allclose(g.rs[:5], [0, step, 2 * step, 3 * step, 4 * step])
    assert np.allclose(g.eta[:5], 5 * np.log(10) / (2 * step) ** 2)
    assert np.allclose(g.eta[5:], 2 * np.log(10) / rc ** 2)
    assert list(g.xi[5:]) == [1, 1, 16, 16] and list(g.lam[5:]) == [1, -1, 1, -1]
    r = g.species_resolved([1, 8])
    assert len(r) == 5 * 2 + 4 * 3
    assert list(zip(r.z1[:4], r.z2[:4])) == [(1, 0), (8, 0), (1, 0), (8, 0)]
    assert list(zip(r.z1[10:13], r.z2[10:13])) == [(1, 1), (1, 8), (8, 8)]


def test_auto_scheme_matches_fixture_table():
    """The auto table + species expansion reproduces a netstat mapping written by the reference
    (restart SD fixture: Auto RCut = 4 A, NRadial = 5, NAngular = 4, species H/O)."""
    from conftest import load_case
    c, _ = load_case("restart__multispecies__globalTargets__optimizer__steepestdesc")
    g = fa.GFunctions.auto(4.0 * fa.AA__BOHR, 5, 4).species_resolved([1, 8])
    assert np.array_equal(g.type, c["in_g_type"])
    assert np.array_equal(g.z1, c["in_g_z1"]) and np.array_equal(g.z2, c["in_g_z2"])
    for k in ("rcut", "eta", "rs", "xi", "lam"):
        assert np.allclose(getattr(g, k), c["in_g_" + k], rtol=1e-7), k
