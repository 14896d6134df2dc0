"""The serving / benchmark step (fnet_run): steady state, HIP graph replays and precision modes.

* After the first step of a batch, fnet_run enqueues every stage without a host read-back (row
  capacities from the previous list); its results equal the first step's bit for bit.
* With fnet_set_graph the steady-state step is captured once and replayed; every replay, with
  eager work and synchronisations in between, equals the eager step bit for bit.
* A list that outgrows the capacities of the previous one (a compressed cell) is detected when
  the results are read and the step is recomputed: results equal a fresh batch's.
* BF16 modes: the split-operand mode holds the gates (energy <= 1e-5 eV/atom, force RMSE
  <= 1e-4 eV/A against the FP64 device path, itself pinned to the oracle); the single-pass mode's
  errors are measured and bounded loosely (they are reported, not gated).
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import workloads as W

pytestmark = pytest.mark.gpu

E_TOL, F_TOL = 1e-5, 1e-4  # eV/atom, eV/A


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.set_graph(False)
    c.close()


def _zscored(ctx, name, scale):
    geos, ns, _ = W.config(name, scale)
    ctx.upload_model(ns)
    b = ctx.batch(geos, fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    return geos, ns


def _results(b):
    Es, Ea = b.energy()
    return Es, Ea, b.forces()


def _same(r1, r2):
    for a, c in zip(r1, r2):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("precision", [fa.FNET_FP64, fa.FNET_FP32, fa.FNET_BF16])
@pytest.mark.parametrize("name,scale", [("water", 0.01), ("si_cells", 0.005), ("si_bulk", 0.25)])
def test_steady_state_and_graph_equal_first_step(ctx, name, scale, precision):
    print("workload", name, scale, "precision", precision, flush=True)
    geos, ns = _zscored(ctx, name, scale)
    ctx.set_graph(False)
    b = ctx.batch(geos, precision)
    b.run()                      # first step: host-sized list
    first = _results(b)
    b.run()                      # steady state, eager
    assert not b.graph_replayed()
    _same(first, _results(b))
    ctx.set_graph(True)
    for rep in range(3):
        b.run()                  # captured on the first call, replayed after
        b.sync()
        _same(first, _results(b))
        assert b.graph_replayed() or rep == 0
        b.acsf()                 # eager work between the replays
    ctx.set_graph(False)
    b.close()


def test_grown_list_is_recomputed(ctx):
    geos, ns = _zscored(ctx, "si64", 0.03)
    ctx.set_graph(True)
    b = ctx.batch(geos, fa.FNET_FP64)
    b.run()
    b.run()
    b.sync()
    coords = np.concatenate([g.coords for g in geos]) * 0.9
    lats = np.stack([g.latvecs for g in geos]) * 0.9
    b.update_coords(coords, lats)   # ~37 % more neighbours per atom: beyond the row capacity
    b.run()
    got = _results(b)
    dense = [fa.Geometry(g.coords * 0.9, g.atnum, g.species, g.latvecs * 0.9) for g in geos]
    ref = ctx.batch(dense, fa.FNET_FP64)
    _same(_results(ref), got)
    b.run()                          # steady state again with the grown capacities
    _same(_results(ref), _results(b))
    ref.close()
    b.close()
    ctx.set_graph(False)


@pytest.mark.parametrize("name,scale", [("si_cells", 0.005), ("si_bulk", 0.25), ("ternary", 0.06),
                                        ("water", 0.01)])
def test_bf16_modes(ctx, name, scale):
    geos, ns = _zscored(ctx, name, scale)
    b = ctx.batch(geos, fa.FNET_FP64)
    Es64, _, F64 = _results(b)
    nat = np.array([g.n for g in geos], float)
    errs = {}
    for mode in (fa.FNET_FP32, fa.FNET_BF16, fa.FNET_BF16_FAST):
        b.set_mode(mode)
        Es, _, F = _results(b)
        e = float((np.abs(Es[:, 0] - Es64[:, 0]) / nat).max())
        f = float(np.sqrt(np.mean((F - F64) ** 2)) / fa.BOHR__AA)
        errs[mode] = (e, f)
    print(name, "fp32 / bf16 / bf16_fast (E eV/atom, F RMSE eV/A):", errs, flush=True)
    for mode in (fa.FNET_FP32, fa.FNET_BF16):
        assert errs[mode][0] <= E_TOL and errs[mode][1] <= F_TOL, (mode, errs[mode])
    # one BF16 pass: 8-bit operands, errors ~1e-3 relative; a wrong operand layout would give O(1)
    e, f = errs[fa.FNET_BF16_FAST]
    assert np.isfinite(e) and np.isfinite(f) and e < 1e-2 and f < 1e-1, (e, f)
    b.close()


TUNES = [dict(fwd_wpb=1), dict(fwd_wpb=4), dict(adj_wpb=2), dict(mlp_waves=1),
         dict(mlp_waves=2), dict(persist=1), dict(fwd_wpb=4, adj_wpb=2, mlp_waves=2, persist=1)]


@pytest.mark.parametrize("precision", [fa.FNET_FP64, fa.FNET_FP32, fa.FNET_BF16])
@pytest.mark.parametrize("name,scale", [("water", 0.01), ("si_cells", 0.005)])
def test_launch_parameters_do_not_change_results(ctx, name, scale, precision):
    """Waves per block, MFMA tile waves and the persistent ACSF grids change how work is
    scheduled, not what each centre or atom computes: energies and forces are bit-identical to
    the default launch, through the first step and the steady state."""
    print("workload", name, scale, "precision", precision, flush=True)
    geos, ns = _zscored(ctx, name, scale)
    ctx.set_graph(False)
    b = ctx.batch(geos, precision)
    b.run()
    ref = _results(b)
    for t in TUNES:
        print("launch parameters", t, flush=True)
        b.set_tune(**t)
        assert {k: v for k, v in b.get_tune().items() if k in t} == t
        b.run()
        _same(ref, _results(b))
        b.run()
        _same(ref, _results(b))
    b.set_tune()
    b.close()


@pytest.mark.parametrize("name,scale", [("water", 0.01), ("si_cells", 0.005)])
def test_autotune_chooses_valid_parameters(ctx, name, scale):
    """The autotuner returns parameters from its candidate sets (or the defaults), the default
    and tuned step times, and leaves results within the FP32 gates of the FP64 path (its radial
    lane split may reorder the FP64 partial sums of the radial features)."""
    print("workload", name, scale, flush=True)
    geos, ns = _zscored(ctx, name, scale)
    b64 = ctx.batch(geos, fa.FNET_FP64)
    E64, _, F64 = _results(b64)
    b64.close()
    b = ctx.batch(geos, fa.FNET_BF16)
    params, t_def, t_tuned = b.autotune(reps=2)
    print("autotune", params, "default %.3f ms tuned %.3f ms" % (t_def, t_tuned))
    # the chosen values, or the defaults (0 / -1) when they did not make the step faster
    assert params["fwd_wpb"] in (0, 1, 2, 4) and params["adj_wpb"] in (0, 1, 2)
    assert params["mlp_waves"] in (0, 1, 2, 4) and params["rbal"] in (-1, 0, 1)
    assert params["persist"] in (0, 1)
    assert t_def > 0 and t_tuned > 0
    assert b.get_tune() == params
    # a second batch of the same model, precision and size class reuses the choice
    b2 = ctx.batch(geos, fa.FNET_BF16)
    assert b2.get_tune() == params
    b2.close()
    b.run()
    Es, _, F = _results(b)
    nat = np.array([g.n for g in geos], float)
    assert (np.abs(Es[:, 0] - E64[:, 0]) / nat).max() <= E_TOL
    assert np.sqrt(np.mean((F - F64) ** 2)) / fa.BOHR__AA <= F_TOL
    b.close()
