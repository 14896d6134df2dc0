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


@pytest.mark.parametrize("kind", ["mse", "mae", "mape", "rms"])
def test_host_batch_loss_equals_device_loss(ctx, kind):
    """The validation monitor's host loss (batch_loss) equals the device loss of
    fnet_train_gradients on the same predictions, targets and weights."""
    from fortnet_amd.train import batch_loss
    c, ds = load_case("restart__multispecies__bothTargets")  # global and atomic targets
    ns = Netstat.from_arrays(c, "in_")
    ctx.upload_model(ns)
    ctx.set_loss(kind)
    geos = dataset.geometries(ds)
    b = ctx.batch(geos, fa.FNET_FP64)
    gt, at = _targets(ds, ns)
    w = ds["weights"].astype(np.float64)
    b.train_gradients(gt, at, ds["atomicweights"], w)
    Es, Ea = b.energy()
    host = batch_loss(Es, Ea, ds["offsets"], ns.n_global, gt, at, ds["atomicweights"], w, kind)
    assert abs(host - b.loss) <= 1e-12 * abs(b.loss)
    b.close()
    ctx.set_loss("mse")


def test_validation_monitor_checkpoints_and_progress(ctx, tmp_path):
    """Five L-BFGS iterations with the training set as validation set (unit datapoint weights,
    so the reference's unweighted validation loss equals the training loss after every
    update), a progress line every 2 iterations and a netstat every 2: the validation losses
    follow the training losses, the lines come at iterations 2, 4, 5, and the netstat written
    at iteration 2 / at the end holds the parameters of a 2-iteration / the full run."""
    from fortnet_amd import h5io
    name = [n for n in SCRATCH if "lbfgs" in n][0]
    c, ds = load_case(name)
    ref = Netstat.from_arrays(c, "after_")
    geos = dataset.geometries(ds)
    gt, at = _targets(ds, ref)
    lines = []
    path = str(tmp_path / "fortnet.hdf5")
    kw = dict(optimizer="lbfgs", learning_rate=float(c["learningrate"]), tol=1e-20,
              min_disp=float(c["mindisp"]),
              max_disp=float(c["maxdisp"]) if float(c["maxdisp"]) > 0 else 1e4,
              memory=int(c["memory"]), line_search=bool(int(c["linemin"])))
    start = Netstat(ref.g, ref.species_z, ref.dims, ref.activation, c["after_params"],
                    ref.zmean, ref.zstd, ref.ext_in, ref.n_global)
    w1 = np.ones(len(geos))
    res = train(ctx, start, geos, gt, at, ds["atomicweights"], w1, n_iter=5, valid_geos=geos,
                valid_g_targets=gt, valid_a_targets=at, valid_atomic_weights=ds["atomicweights"],
                n_print=2, n_save=2, netstat_path=path, out=lines.append, **kw)
    print("losses", res.losses, "valid", res.valid_losses)
    assert len(res.valid_losses) == res.iterations == 5
    assert np.allclose(res.valid_losses, res.train_losses, rtol=1e-12, atol=0)
    assert [int(l.split()[0]) for l in lines] == [2, 4, 5]
    assert all(len(l.split()) == 4 for l in lines)
    final = h5io.load_netstat(path)
    assert np.array_equal(final.params, res.params)
    res2 = train(ctx, start, geos, gt, at, ds["atomicweights"], w1, n_iter=2, n_save=2,
                 netstat_path=str(tmp_path / "two.hdf5"), **kw)
    assert np.array_equal(h5io.load_netstat(str(tmp_path / "two.hdf5")).params, res2.params)
