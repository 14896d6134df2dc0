"""Continuous-batching serving: each result is the same alone and in the mix.

Requests are served by the batching loop (several structures per device batch, two batches in
flight); every request is then evaluated alone in its own batch.  Per-atom work is the same in
both, so energies agree to FP64 rounding of the per-structure sum order (identical order here)
and forces to the order of the FP64 atomic accumulation: tolerance 1e-10 relative.
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import serve
from fortnet_amd import workloads as W

pytestmark = pytest.mark.gpu


def _model_and_pool(n=10):
    geos, ns, _ = W.config("ternary", 0.2)
    geos = sorted(geos, key=lambda g: g.n)[:n]
    ctx = fa.Context(0)
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.close()
    return ns, geos


@pytest.mark.parametrize("precision", [fa.FNET_FP32, fa.FNET_FP64], ids=["fp32", "fp64"])
def test_results_same_alone_and_in_the_mix(precision):
    ns, pool = _model_and_pool()
    print("pool sizes", [g.n for g in pool], flush=True)
    arrivals = np.concatenate([np.zeros(6), 0.002 * np.arange(1, 9)])
    reqs = [serve.Request(k, pool[k % len(pool)], float(t)) for k, t in enumerate(arrivals)]
    srv = serve.BatchingServer(ns, precision, max_atoms_per_batch=4000)
    try:
        srv.run(reqs)
    finally:
        srv.close()
    assert max(srv.batch_sizes) > 1, srv.batch_sizes
    ctx = fa.Context(0)
    ctx.upload_model(ns)
    for r in reqs:
        b = ctx.batch([r.geo], precision)
        E, _ = b.energy()
        F = b.forces()
        b.close()
        assert abs(E[0, 0] - r.energy[0]) <= 1e-10 * max(1.0, abs(E[0, 0])), r.rid
        assert np.abs(F - r.forces).max() <= 1e-10 * max(1.0, np.abs(F).max()), r.rid
        assert r.done >= r.arrival
    ctx.close()


def test_serial_loop_matches_batching():
    ns, pool = _model_and_pool(4)
    reqs_a = [serve.Request(k, pool[k], 0.0) for k in range(4)]
    reqs_b = [serve.Request(k, pool[k], 0.0) for k in range(4)]
    serve.serial_loop(ns, reqs_a)
    srv = serve.BatchingServer(ns)
    try:
        srv.run(reqs_b)
    finally:
        srv.close()
    for a, b in zip(reqs_a, reqs_b):
        assert abs(a.energy[0] - b.energy[0]) <= 1e-10 * max(1.0, abs(a.energy[0]))
        assert np.abs(a.forces - b.forces).max() <= 1e-10 * max(1.0, np.abs(a.forces).max())
    s = serve.latency_summary(reqs_b)
    assert s["p99_ms"] >= s["p50_ms"] > 0
