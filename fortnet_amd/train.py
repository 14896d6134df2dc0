"""Training loop: initialisation, z-score, device gradients, host optimizers.

Mirrors ``TBpnn_nTrain`` (``bpnn.F90:147-395``): the gradient of the whole training set is
computed at the initial parameters, then each iteration hands the loss and the normalised
gradient of every element network to that network's optimizer, uploads the new parameters,
re-draws the datapoint shuffle when requested and recomputes loss and gradient.  Gradients,
losses and the regularisation run on the GPU (``fnet_train_gradients``); the optimizers are the
host state machines of :mod:`fortnet_amd.optim`.  The shuffle only permutes the order in which
the full-batch gradient is summed, so it changes no result beyond rounding; its draws are still
taken from the generator so the stream stays aligned with the reference's.
"""
import numpy as np

from . import optim
from ._lib import FNET_FP64
from .core import Netstat
from .ranlux import Ranlux, initial_params, knuth_shuffle


def zscore_stats(G, offsets, dp_weights=None):
    """Means and standard deviations of the features over the training atoms, weighted by
    the datapoint weights (``TAcsf_getMeansAndVariances``, acsf.F90:445-490)."""
    S = len(offsets) - 1
    w = np.ones(S) if dp_weights is None else np.asarray(dp_weights, np.float64)
    wa = np.repeat(w, np.diff(offsets))
    n = wa.sum()
    mean = (wa[:, None] * G).sum(axis=0) / n
    std = np.sqrt((wa[:, None] * (G - mean) ** 2).sum(axis=0) / n)
    return mean, std


class TrainResult:
    def __init__(self, params, losses, grad_norms, iterations, converged):
        self.params = params
        self.losses = losses
        self.grad_norms = grad_norms
        self.iterations = iterations
        self.converged = converged


def train(ctx, netstat, geos, g_targets=None, a_targets=None, atomic_weights=None,
          dp_weights=None, optimizer="sd", n_iter=1, learning_rate=0.01, tol=1e-20,
          min_disp=1e-6, max_disp=1e4, memory=1000, line_search=True, loss="mse",
          regularization=None, reg_strength=0.0, reg_alpha=0.0, shuffle=False, rng=None,
          precision=FNET_FP64, device_optimizer=False):
    """Runs ``n_iter`` training iterations from the parameters of ``netstat`` (uploaded as the
    context's model) and returns a :class:`TrainResult`; the context keeps the final model.
    ``device_optimizer``: the FP64 optimizer step runs on the GPU (``fnet_optimizer_step``)
    instead of the host state machines."""
    ctx.upload_model(netstat)
    ctx.set_loss(loss)
    ctx.set_regularization(regularization, reg_strength, reg_alpha)
    b = ctx.batch(geos, precision)
    nsp, npar = netstat.params.shape
    opts = []
    if device_optimizer:
        ctx.optimizer_init(optimizer, tol, min_disp, max_disp, learning_rate, memory, line_search)
    else:
        for sp in range(nsp):
            o = optim.make(optimizer, npar, tol, min_disp, max_disp, learning_rate, memory,
                           line_search)
            o.reset(netstat.params[sp])
            opts.append(o)
    order = list(range(len(geos)))
    if shuffle and rng is not None:
        knuth_shuffle(rng, order)
    grad, _ = b.train_gradients(g_targets, a_targets, atomic_weights, dp_weights)
    loss_val = b.loss
    params = netstat.params.copy()
    losses, norms = [loss_val], []
    converged = False
    it = 0
    nw = int(sum(int(netstat.dims[l]) * int(netstat.dims[l + 1])
                 for l in range(len(netstat.dims) - 1)) + int(netstat.dims[-1]))
    for it in range(1, n_iter + 1):
        norms.append(float(np.sqrt(np.sum(grad ** 2))))
        if device_optimizer:
            params, converged = ctx.optimizer_step()
        else:
            new = np.empty_like(params)
            for sp in range(nsp):
                fx = loss_val + _regu_loss(params[sp, :nw], regularization, reg_strength,
                                           reg_alpha)
                new[sp], converged = opts[sp].next(fx, grad[sp])
            params = new
            ctx.upload_params(params)
        if shuffle and rng is not None:
            knuth_shuffle(rng, order)
        grad, _ = b.train_gradients(g_targets, a_targets, atomic_weights, dp_weights)
        loss_val = b.loss
        losses.append(loss_val)
        if converged:
            break
    b.close()
    return TrainResult(params, np.array(losses), np.array(norms), it, converged)


def _regu_loss(w, kind, strength, alpha):
    """Regularisation loss of one network's weights (loss.F90:119-196)."""
    if not kind or strength == 0.0:
        return 0.0
    n = float(len(w))
    if kind == "lasso":
        return strength / n * float(np.sum(np.abs(w)))
    if kind == "ridge":
        return strength / (2.0 * n) * float(np.sum(w ** 2))
    return strength / n * ((1.0 - alpha) / 2.0 * float(np.sum(w ** 2))
                           + alpha * float(np.sum(np.abs(w))))


def scratch_netstat(gfuncs, species_z, dims, activation, seed, zmean=None, zstd=None,
                    ext_in=None, n_global=None):
    """A netstat with RANLUX/Xavier initial weights (``TBpnn_init``); returns it together with
    the generator, whose stream continues with the training loop's shuffles."""
    params, gen = initial_params(dims, len(species_z), seed)
    ns = Netstat(gfuncs, species_z, dims, activation, params, zmean, zstd, ext_in, n_global)
    return ns, gen
