"""Continuous-batching serving: each result is the same alone and in the mix, and equals the oracle.

Requests are served by the batching loop (several structures per device batch, two batches in
flight); every request is then evaluated alone in its own batch.  Per-atom work is the same in
both (forces are gathered per atom in a fixed order), so results agree to 1e-10 relative.  Every
served result is also held against the FP64 oracle: FP64 at 1e-10 relative, FP32 within the
accuracy gates (energy 1e-5 eV/atom, force RMSE 1e-4 eV/A).
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import serve
from fortnet_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_energy_forces(ns, geo):
    """FP64 oracle energy (sum of the atomic outputs) and analytic forces of one structure."""
    gf = O.GFunctions(ns.g.type, ns.g.rcut, ns.g.eta, ns.g.rs, ns.g.kappa, ns.g.lam, ns.g.xi,
                      ns.g.atomid, ns.g.z1, ns.g.z2)
    nets = O.Nets(ns.dims, ns.activation, ns.params)
    st = O.Structure(geo.coords, geo.latvecs, int(geo.periodic), geo.atnum)
    X = O.zscore(O.acsf(st, gf), ns.zmean, ns.zstd)
    E = nets.predict(X, geo.species).sum()
    J = nets.jacobian(X, geo.species)
    F = O.acsf_adjoint(st, gf, J[:, 0, :] * O.zscale(ns.zstd, gf.n)[None, :])
    return E, F


def check_against_oracle(ns, geo, energy, forces, precision):
    E0, F0 = oracle_energy_forces(ns, geo)
    if precision == fa.FNET_FP64:
        assert abs(energy - E0) <= 1e-10 * max(1.0, abs(E0))
        assert np.abs(forces - F0).max() <= 1e-9 * max(1.0, np.abs(F0).max())
    else:
        assert abs(energy - E0) / geo.n <= 1e-5
        assert np.sqrt(np.mean((forces - F0) ** 2)) / fa.BOHR__AA <= 1e-4


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
    for r in reqs[:len(pool)]:
        check_against_oracle(ns, r.geo, r.energy[0], r.forces, precision)


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
