"""Host-side training pieces on the CPU: optimizers, RANLUX stream, Xavier initialisation.

The optimizers are checked against what the mathematics fixes on quadratic objectives (exact
steepest-descent step and its displacement cap, exact line minimum, convergence of CG / L-BFGS /
FIRE to the minimiser); the generator against its structure (24-bit words, reproducible stream
however it is split, uniform moments); the initialisation against its closed form.  The values
the reference produces are pinned on the GPU by the training fixtures (test_gpu_training.py).
"""
import math

import numpy as np
import pytest

from fortnet_amd import optim
from fortnet_amd.ranlux import Ranlux, initial_params, knuth_shuffle, xavier_weights
from fortnet_amd.core import nparams


def _quadratic(n, seed=0):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, n))
    A = q @ q.T + n * np.eye(n)
    xs = rng.normal(size=n)
    f = lambda x: 0.5 * (x - xs) @ A @ (x - xs)
    g = lambda x: A @ (x - xs)
    return f, g, xs


def test_sd_step_and_cap():
    f, g, _ = _quadratic(5)
    x0 = np.zeros(5)
    o = optim.SteepestDescent(5, 1e-12, 1e4, 0.01)
    o.reset(x0)
    x1, conv = o.next(f(x0), g(x0))
    assert not conv and np.allclose(x1, x0 - 0.01 * g(x0), rtol=0, atol=1e-15)
    o = optim.SteepestDescent(5, 1e-12, 1e-3, 0.01)
    o.reset(x0)
    x1, _ = o.next(f(x0), g(x0))
    step = x1 - x0
    assert abs(np.abs(step).max() - 1e-3) < 1e-15
    assert np.allclose(step / np.linalg.norm(step), -g(x0) / np.linalg.norm(g(x0)))


def test_linemin_finds_exact_minimum_of_quadratic():
    f, g, xs = _quadratic(4, 1)
    x0 = np.zeros(4)
    d = -g(x0)
    lm = optim.LineMin(4, 10, 1e-12, 1e4)
    lm.reset(x0, d, 1.0)
    x = x0
    for _ in range(10):
        x, conv = lm.next(f(x), g(x))
        if conv:
            break
    assert conv
    dn = d / np.linalg.norm(d)
    A = np.array([g(np.eye(4)[k]) - g(np.zeros(4)) for k in range(4)]).T
    lam = -(g(x0) @ dn) / (dn @ A @ dn)
    assert np.allclose(lm.min_x(), x0 + lam * dn, atol=1e-10)


@pytest.mark.parametrize("kind", ["cg", "lbfgs", "fire"])
def test_optimizers_converge_on_quadratic(kind):
    n = 6
    f, g, xs = _quadratic(n, 2)
    o = optim.make(kind, n, 1e-9, 1e-6, 1.0, memory=5)
    x = np.zeros(n)
    o.reset(x)
    for _ in range(5000):
        x, conv = o.next(f(x), g(x))
        if conv:
            break
    assert np.abs(x - xs).max() < 1e-6


def test_first_steps_closed_forms():
    """One call from x0: CG moves min(5|g|, cap) along -g/|g|, L-BFGS min(1, cap) along it,
    FIRE -(0.1 dt_max)^2 g (conjgrad.F90, lbfgs.F90 + linemin.F90, fire.F90)."""
    n = 5
    f, g, _ = _quadratic(n, 3)
    x0 = np.ones(n)
    gv = g(x0)
    u = -gv / np.linalg.norm(gv)
    for kind, lam in (("cg", 5 * np.linalg.norm(gv)), ("lbfgs", 1.0)):
        for cap in (1e4, 1e-2):
            o = optim.make(kind, n, 1e-20, 1e-10, cap)
            o.reset(x0)
            x1, _ = o.next(f(x0), gv)
            step = min(lam, cap / np.abs(u).max())
            assert np.allclose(x1, x0 + step * u, rtol=1e-14, atol=1e-14), kind
    o = optim.make("fire", n, 1e-20, 1e-10, 0.1)
    o.reset(x0)
    x1, _ = o.next(f(x0), gv)
    assert np.allclose(x1, x0 - 0.01 ** 2 * gv, rtol=0, atol=1e-15)


def test_ranlux_stream_structure():
    a = Ranlux(123456).uniform(1000)
    r = Ranlux(123456)
    b = np.concatenate([r.uniform(7), r.uniform(1), r.uniform(992)])
    assert np.array_equal(a, b)  # the stream does not depend on how it is drawn
    assert np.all((a > 0.0) & (a < 1.0))
    # 24-bit resolution: u * 2^24 is integral unless u < 2^-12 (then 24 more bits)
    big = a[a >= 2.0 ** -12]
    assert np.all(big * 2 ** 24 == np.floor(big * 2 ** 24))
    c = Ranlux(123456).uniform(200000)
    assert abs(c.mean() - 0.5) < 0.005 and abs(c.var() - 1.0 / 12.0) < 0.002
    assert not np.array_equal(Ranlux(1).uniform(50), Ranlux(2).uniform(50))


def test_xavier_closed_form_and_layout():
    r = Ranlux(7)
    w = xavier_weights(r, 3, 5)
    u = Ranlux(7).uniform(15)
    var = 2.0 / 8.0
    bound = math.sqrt(-2.0 * var * math.log(0.1 * math.sqrt(2.0 * math.pi * var)))
    x = (u - 0.5) * 2.0 * bound
    want = np.exp(-0.5 * x ** 2 / var) / math.sqrt(2.0 * math.pi * var)
    assert np.allclose(w, want.reshape(5, 3).T, rtol=0, atol=1e-16)  # ww(3, 5), column-major
    assert np.all(w >= 0.1 - 1e-12)  # the truncation: every weight is at least min = 0.1
    dims = [4, 3, 2, 1]
    p, _ = initial_params(dims, 2, 123456)
    assert p.shape == (2, nparams(dims))
    nw = 4 * 3 + 3 * 2 + 2 * 1
    assert np.all(p[:, :nw] > 0) and np.all(p[:, nw:] == 0.0)  # dummy layer and biases zero


def test_knuth_shuffle_is_a_permutation():
    order = knuth_shuffle(Ranlux(5), list(range(50)))
    assert sorted(order) == list(range(50)) and order != list(range(50))
