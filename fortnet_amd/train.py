"""Training loop: initialisation, z-score, device gradients, host optimizers.

Mirrors ``TBpnn_nTrain`` (``bpnn.F90:147-395``): the gradient of the whole training set is
computed at the initial parameters, then each iteration hands the loss and the normalised
gradient of every element network to that network's optimizer, uploads the new parameters,
re-draws the datapoint shuffle when requested and recomputes loss and gradient.  Gradients,
losses and the regularisation run on the GPU (``fnet_train_gradients``); the optimizers are the
host state machines of :mod:`fortnet_amd.optim` (or the device ones).  The shuffle only permutes
the order in which the full-batch gradient is summed, so it changes no result beyond rounding;
its draws are still taken from the generator so the stream stays aligned with the reference's.

As in the reference, each iteration records the training loss and the gradient norm, and with a
validation set the loss of its predictions (``tMonitorValid``, bpnn.F90:283-291,319-327); it
prints progress every ``n_print`` iterations (bpnn.F90:329-338) and writes the netstat every
``n_save`` iterations, at the last one and on convergence (bpnn.F90:340-359).  ``write_iterout``
writes the trajectory file of ``writeIterTrajToFile`` (iterout.F90).
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
    """``losses``: the loss at the initial parameters, then after every iteration;
    ``grad_norms``: the norm of the gradient each iteration's step used; ``valid_losses``: the
    validation loss after every iteration (None without a validation set).  The reference's
    trajectory (trainLoss, gradients, validLoss of TBpnn_nTrain) is ``losses[1:]``,
    ``grad_norms``, ``valid_losses``."""

    def __init__(self, params, losses, grad_norms, iterations, converged, valid_losses=None):
        self.params = params
        self.losses = losses
        self.grad_norms = grad_norms
        self.iterations = iterations
        self.converged = converged
        self.valid_losses = valid_losses

    @property
    def train_losses(self):
        return self.losses[1:]


def _item_loss(kind, d, t):
    """Loss of each row of residuals d = t - p over its targets t (loss.F90:285-366)."""
    n = d.shape[1]
    if kind == "mae":
        return np.abs(d).sum(axis=1) / n
    if kind == "mape":
        return 100.0 * np.abs(d / t).sum(axis=1) / n
    if kind == "rms":
        return np.sqrt((d * d).sum(axis=1) / n)
    return (d * d).sum(axis=1) / n


def batch_loss(E_struct, E_atom, offsets, n_global, g_targets=None, a_targets=None,
               atomic_weights=None, dp_weights=None, kind="mse"):
    """Loss of a set of predictions (maLoss / mapLoss / msLoss / rmsLoss, loss.F90:370-700), in
    FP64 on the host: global targets weighted by w_s times the structure's summed atomic weights,
    atomic targets by w_s times the atom's weight, divided by the summed weights."""
    offsets = np.asarray(offsets)
    S = len(offsets) - 1
    N = int(offsets[-1])
    w = np.ones(S) if dp_weights is None else np.asarray(dp_weights, np.float64)
    aw = np.ones(N) if atomic_weights is None else np.asarray(atomic_weights, np.float64)
    sid = np.repeat(np.arange(S), np.diff(offsets))
    num = den = 0.0
    nout = E_atom.shape[1]
    if n_global < nout:
        at = np.asarray(a_targets, np.float64).reshape(N, nout - n_global)
        li = _item_loss(kind, at - E_atom[:, n_global:], at)
        num += float(np.sum(w[sid] * aw * li))
        den += float(np.sum(w[sid] * aw))
    if n_global > 0:
        gt = np.asarray(g_targets, np.float64).reshape(S, n_global)
        saw = np.bincount(sid, weights=aw, minlength=S)
        ls = _item_loss(kind, gt - E_struct[:, :n_global], gt)
        num += float(np.sum(w * saw * ls))
        den += float(np.sum(w * saw))
    return num / den


def _fortran_e(x, w, d):
    """Fortran's Ew.d edit descriptor: 0.ddd...E+ee right-aligned in w columns (exponents beyond
    two digits drop the E, as Fortran writes them)."""
    if x == 0.0 or not np.isfinite(x):
        m, e = 0.0, 0
    else:
        e = int(np.floor(np.log10(abs(x)))) + 1
        m = x / 10.0 ** e
        if round(abs(m), d) >= 1.0:
            m /= 10.0
            e += 1
    digits = ("%.*f" % (d, abs(m)))[1:]  # ".ddd"
    sign = "+" if e >= 0 else "-"
    exp = ("E%s%02d" % (sign, abs(e))) if abs(e) < 100 else ("%s%03d" % (sign, abs(e)))
    return ("%s0%s%s" % ("-" if m < 0 else "", digits, exp)).rjust(w)


def progress_line(it, loss, grad_norm, valid_loss=None):
    """A progress line of the training loop, format (I10,5X,E15.6,4X,E15.6[,4X,E15.6])
    (bpnn.F90:329-338)."""
    out = "%10d%s%s%s%s" % (it, " " * 5, _fortran_e(loss, 15, 6), " " * 4,
                            _fortran_e(grad_norm, 15, 6))
    if valid_loss is not None:
        out += " " * 4 + _fortran_e(valid_loss, 15, 6)
    return out


def _es(x, w=26, d=16):
    """Fortran's ESw.dE3 edit descriptor."""
    m, e = ("%.*E" % (d, x)).split("E")
    e = int(e)
    return ("%sE%s%03d" % (m, "+" if e >= 0 else "-", abs(e))).rjust(w)


def write_iterout(path, train_losses=None, grad_norms=None, valid_losses=None):
    """The iteration trajectory file ``iterout.dat`` of ``writeIterTrajToFile``
    (iterout.F90): per iteration its number (width of the largest number) and the training loss,
    the gradient norm and the validation loss that are given, each as ES26.16E3."""
    cols = [np.asarray(c, np.float64) for c in (train_losses, grad_norms, valid_losses)
            if c is not None]
    if not cols:
        raise ValueError("no trajectory to write")
    n = len(cols[0])
    w = len(str(n))
    with open(path, "w") as fh:
        for k in range(n):
            fh.write("%*d%s\n" % (w, k + 1, "".join(_es(float(c[k])) for c in cols)))


def train(ctx, netstat, geos, g_targets=None, a_targets=None, atomic_weights=None,
          dp_weights=None, optimizer="sd", n_iter=1, learning_rate=0.01, tol=1e-20,
          min_disp=1e-6, max_disp=1e4, memory=1000, line_search=True, loss="mse",
          regularization=None, reg_strength=0.0, reg_alpha=0.0, shuffle=False, rng=None,
          precision=FNET_FP64, device_optimizer=False, valid_geos=None, valid_g_targets=None,
          valid_a_targets=None, valid_atomic_weights=None, n_print=None, n_save=None,
          netstat_path=None, out=print):
    """Runs ``n_iter`` training iterations from the parameters of ``netstat`` (uploaded as the
    context's model) and returns a :class:`TrainResult`; the context keeps the final model.
    ``device_optimizer``: the FP64 optimizer step runs on the GPU (``fnet_optimizer_step``)
    instead of the host state machines.  ``valid_geos`` (+ targets and atomic weights): the
    validation set, whose loss is computed after every iteration (without datapoint weights, as
    the reference).  ``n_print``: a progress line through ``out`` every n_print iterations and
    at the last; ``n_save`` with ``netstat_path``: the netstat (HDF5) written every n_save
    iterations, at the last one and on convergence."""
    ctx.upload_model(netstat)
    ctx.set_loss(loss)
    ctx.set_regularization(regularization, reg_strength, reg_alpha)
    b = ctx.batch(geos, precision)
    bv = None
    valid_losses = None
    if valid_geos is not None:
        bv = ctx.batch(valid_geos, precision)
        voff = np.concatenate([[0], np.cumsum([g.n for g in valid_geos])])
        valid_losses = []

    def valid_loss():
        Es, Ea = bv.energy()
        return batch_loss(Es, Ea, voff, netstat.n_global, valid_g_targets, valid_a_targets,
                          valid_atomic_weights, None, loss)

    def save(p):
        from . import h5io
        h5io.write_netstat(netstat_path, Netstat(netstat.g, netstat.species_z, netstat.dims,
                                                 netstat.activation, p, netstat.zmean,
                                                 netstat.zstd, netstat.ext_in, netstat.n_global))
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
        if bv is not None:
            valid_losses.append(valid_loss())
        vl = valid_losses[-1] if bv is not None else None
        if n_print and (it % n_print == 0 or it == n_iter):
            out(progress_line(it, loss_val, norms[-1], vl))
        if netstat_path and n_save and (it % n_save == 0 or it == n_iter):
            save(params)
        if converged:
            if n_print:
                out(progress_line(it, loss_val, norms[-1], vl))
            if netstat_path:
                save(params)
            break
    b.close()
    if bv is not None:
        bv.close()
        valid_losses = np.array(valid_losses)
    return TrainResult(params, np.array(losses), np.array(norms), it, converged, valid_losses)


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
