"""The no-allocation force assembly (fnet_forces_noalloc) against fnet_forces, through the C-ABI.

fnet_forces_noalloc runs the same kernels as fnet_forces with its per-entry scratch in the static
device array: on two calls in a row (the second on a smaller batch) the forces equal those of
fnet_forces bit for bit, in both launch modes and precisions; a list beyond the scratch's
32 MiB is refused before anything is launched.
"""
import ctypes

import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import _lib
from fortnet_amd import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(ctx, b):
    lib = _lib.load()
    _lib.check(ctx.h, lib.fnet_forces_noalloc(ctx.h, b.h), "fnet_forces_noalloc")


def _zscored(ctx, name, scale):
    geos, ns, _ = W.config(name, scale)
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    return geos


@pytest.mark.parametrize("launch", [fa.FNET_BATCHED, fa.FNET_PER_STRUCT], ids=["batched", "per_struct"])
@pytest.mark.parametrize("precision", [fa.FNET_FP64, fa.FNET_FP32], ids=["fp64", "fp32"])
def test_noalloc_forces_equal_fnet_forces(precision, launch):
    ctx = fa.Context(0)
    for name, scale in (("water", 0.01), ("si64", 0.02)):  # the second call on a smaller batch
        print("workload", name, scale, "precision", precision, "launch", launch, flush=True)
        geos = _zscored(ctx, name, scale)
        b = ctx.batch(geos, precision, launch)
        b.run()
        b.run()
        F_ref = b.forces()
        _noalloc(ctx, b)
        F = b.forces()
        assert np.array_equal(F, F_ref)
        b.close()
    ctx.close()


def test_noalloc_refuses_large_lists():
    ctx = fa.Context(0)
    geos = _zscored(ctx, "water", 0.2)  # ~2 M list entries: beyond the 32 MiB scratch
    b = ctx.batch(geos, fa.FNET_FP32)
    F_ref = b.forces()
    with pytest.raises(fa.FnetError, match="too large"):
        _noalloc(ctx, b)
    assert np.array_equal(b.forces(), F_ref)
    b.close()
    ctx.close()
