"""GPU parity of the training path against the reference's training fixtures.

The reference's ``training/*`` and ``input/*`` regression cases train for one iteration and store
the network after it (``_fortnet.hdf5``).  75 of them start from scratch: RANLUX (seed 123456,
luxury 3) / Xavier initial weights, the z-score of the training set, one step of SD, CG,
L-BFGS or FIRE with the mse, mae, mape or rms loss and lasso / ridge / elastic-net
regularisation.  Reproducing the stored parameters pins, in one chain, the generator and the
initialisation (random.F90, ranlux.F90, network.F90:95-122), the z-score statistics
(acsf.F90:445-490), the activation of every case's network, the loss gradients (loss.F90), the
regularisation (nestedtypes.F90:336-370), the gradient normalisation (bpnn.F90:610-766) and each
optimizer's first step.  Parameters must agree element-wise to the reference's tolerance,
``np.allclose(atol=1e-10, rtol=1e-9)`` (testwithworkdir.py:24-25,301-525), in FP64.
The restart-style cases of ``input/*`` (an input netstat, one SD step) are checked the same way.
Every case runs twice: with the optimizer step on the host state machines and with the FP64
device optimizers (``fnet_optimizer_step``).
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import dataset
from fortnet_amd.core import Netstat
from fortnet_amd.train import scratch_netstat, train, zscore_stats
from conftest import case_names, load_case

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-10, 1e-9


def _cases(scratch):
    out = []
    for n in case_names():
        if not n.startswith(("input", "training")):
            continue
        c, _ = load_case(n)
        if str(c["mode"]) != "train" or "after_params" not in c:
            continue
        if int(c["scratch"]) == scratch and (scratch or str(c["optimizer"]) == "sd"):
            out.append(n)
    return out


SCRATCH = _cases(1)
RESTART = _cases(0)


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def _targets(ds, ns):
    ng = ns.n_global
    gt = ds["globaltargets"] if ng > 0 else None
    at = ds["atomictargets"] if ns.n_out - ng > 0 else None
    return gt, at


def _close(got, want, what):
    # element-wise, as the reference's netstat comparison (testwithworkdir.py:301-525)
    err = float(np.abs(got - want).max())
    assert np.allclose(got, want, atol=ATOL, rtol=RTOL), "%s: max abs err %.3e" % (what, err)


def _run(ctx, c, ds, ns, rng, device_optimizer=False):
    gt, at = _targets(ds, ns)
    regu = str(c["regu_type"]) or None
    return train(ctx, ns, dataset.geometries(ds), gt, at, ds["atomicweights"],
                 ds["weights"].astype(np.float64), optimizer=str(c["optimizer"]),
                 n_iter=int(c["niterations"]), learning_rate=float(c["learningrate"]),
                 tol=float(c["threshold"]), min_disp=float(c["mindisp"]),
                 max_disp=float(c["maxdisp"]) if float(c["maxdisp"]) > 0 else 1e4,
                 memory=int(c["memory"]), line_search=bool(int(c["linemin"])),
                 loss=str(c["loss"]), regularization=regu,
                 reg_strength=float(c["regu_strength"]), reg_alpha=float(c["regu_alpha"]),
                 shuffle=str(c["shuffle"]) == "yes", rng=rng, device_optimizer=device_optimizer)


# the optimizer step on the host state machines or on the device (fnet_optimizer_step)
OPT_WHERE = [False, True]
OPT_IDS = ["host_opt", "device_opt"]


@pytest.mark.parametrize("device_opt", OPT_WHERE, ids=OPT_IDS)
@pytest.mark.parametrize("name", SCRATCH)
def test_training_from_scratch(ctx, name, device_opt):
    c, ds = load_case(name)
    print("case", name, "device optimizer", device_opt, flush=True)
    ref = Netstat.from_arrays(c, "after_")
    geos = dataset.geometries(ds)
    zmean = zstd = None
    if "after_zmean" in c:
        # the z-score of the training set, computed here from the FP64 ACSF
        raw = Netstat(ref.g, ref.species_z, ref.dims, ref.activation, ref.params, None, None,
                      ref.ext_in, ref.n_global)
        ctx.upload_model(raw)
        b = ctx.batch(geos, fa.FNET_FP64)
        G = b.acsf()
        b.close()
        zmean, zstd = zscore_stats(G, ds["offsets"], ds["weights"].astype(np.float64))
        _close(zmean, c["after_zmean"], "z-score means")
        _close(zstd, c["after_zstd"], "z-score deviations")
    ns, gen = scratch_netstat(ref.g, ref.species_z, ref.dims, ref.activation, int(c["seed"]),
                              zmean, zstd, ref.ext_in, ref.n_global)
    res = _run(ctx, c, ds, ns, gen, device_opt)
    _close(res.params, c["after_params"], "parameters after one iteration")


@pytest.mark.parametrize("device_opt", OPT_WHERE, ids=OPT_IDS)
@pytest.mark.parametrize("name", RESTART)
def test_training_restart(ctx, name, device_opt):
    c, ds = load_case(name)
    print("case", name, "device optimizer", device_opt, flush=True)
    ns = Netstat.from_arrays(c, "in_")
    res = _run(ctx, c, ds, ns, None, device_opt)
    _close(res.params, c["after_params"], "parameters after one iteration")
