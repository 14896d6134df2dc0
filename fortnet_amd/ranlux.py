"""RANLUX luxury pseudorandom numbers and the network initialisation that draws them.

Fortnet seeds one generator per run (``initprogram.F90:392``: luxury level 3, the ``RandomSeed``
of the input) and draws, in order: the initial weights of every element network
(``bpnn.F90:136`` -> ``network.F90:95-122`` -> ``layer.F90:58-88``), then the datapoint shuffles
of the training loop (``bpnn.F90:270,310``).  The generator is James's RANLUX (F. James, Comput.
Phys. Commun. 79 (1994) 111; M. Luscher, ibid. 100): a subtract-with-borrow generator with
24-bit words, lags 24 and 10, that discards ``nskip`` numbers after every 24 delivered
(``ranlux.F90:92-390``).  The seeding, the 24-bit arithmetic and the delivery rule are the
published algorithm; they are restated here because the initial weights must match the
reference's bit for bit.
"""
import math

import numpy as np

_NSKIP = (0, 24, 73, 199, 365)  # numbers discarded per 24 delivered, luxury levels 0..4
_TWO24 = 1 << 24
_MASK24 = _TWO24 - 1
_ICONS = 2147483563


class Ranlux:
    """RANLUX generator (default luxury level 3, default seed 314159265)."""

    def __init__(self, seed=314159265, luxury=3):
        if not 0 <= luxury <= 4:
            luxury = 3
        self.nskip = _NSKIP[luxury]
        jseed = seed if seed > 0 else 314159265
        self.seeds = [0] * 24
        # seeds from a multiplicative congruential generator (Park-Miller with Schrage's method)
        for i in range(24):
            k = jseed // 53668
            jseed = 40014 * (jseed - k * 53668) - k * 12211
            if jseed < 0:
                jseed += _ICONS
            self.seeds[i] = jseed % _TWO24
        self.carry = 1 if self.seeds[23] & ~_MASK24 else 0
        # lag pointers (0-based): i walks 23, 22, ..., 0, 23, ...; j starts 14 behind
        self.i = 23
        self.j = 9
        self.n24 = 0
        self.twom24 = 2.0 ** -24
        self.twom12 = 2.0 ** -12

    def _step(self):
        s = self.seeds
        v = s[self.j] - s[self.i] - self.carry
        if v & ~_MASK24:
            v &= _MASK24
            self.carry = 1
        else:
            self.carry = 0
        s[self.i] = v
        self.i = self.i - 1 if self.i > 0 else 23
        self.j = self.j - 1 if self.j > 0 else 23
        return v

    def uniform(self, n):
        """n numbers in (0, 1) (``getRandom``); small values get 24 further bits."""
        out = np.empty(n)
        for q in range(n):
            v = self._step()
            u = v * self.twom24
            if u < self.twom12:
                u += (self.seeds[self.j] * self.twom24) * self.twom24
                if u <= 0.0:
                    u = self.twom24 * self.twom24
            out[q] = u
            self.n24 += 1
            if self.n24 == 24:
                self.n24 = 0
                for _ in range(self.nskip):
                    self._step()
        return out


def knuth_shuffle(gen, order):
    """In-place Knuth shuffle of a list of datapoint indices (``random.F90:39-57``)."""
    n = len(order)
    for i in range(n, 1, -1):
        r = gen.uniform(1)[0]
        p = int(r * i)  # 0-based position of the element swapped with order[i - 1]
        order[p], order[i - 1] = order[i - 1], order[p]
    return order


def xavier_weights(gen, n_current, n_next, gain=1.0, minval=0.1):
    """``normalXavier2d`` (``random.F90:122-181``) for a layer ``ww(n_current, n_next)``: uniform
    draws in column-major order, stretched to [-b, b] with b from the minimum value, mapped
    through the normal density of variance 2 gain^2 / (n_current + n_next).  Returns the matrix
    as ``ww[k, o]``."""
    var = gain * gain * 2.0 / float(n_next + n_current)
    bound = math.sqrt(-2.0 * var * math.log(minval * math.sqrt(2.0 * math.pi * var)))
    r = gen.uniform(n_current * n_next)
    x = (r - 0.5) * 2.0 * bound
    w = np.exp(-0.5 * x ** 2 / var) / math.sqrt(2.0 * math.pi * var)
    return w.reshape(n_next, n_current).T  # column-major fill of ww(n_current, n_next)


def initial_params(dims, n_species, seed, luxury=3, gen=None):
    """Initial serialised parameters [n_species, nparams] of ``TBpnn_init`` (``bpnn.F90:136``):
    per species, per layer in order, Xavier draws for ``ww(d_l, d_{l+1})``; the dummy output
    layer ``ww(d_n, 1)`` is drawn and then zeroed, and all biases are zero
    (``network.F90:112-120``).  Returns (params, generator) so the caller can continue the
    stream (datapoint shuffles)."""
    dims = [int(d) for d in dims]
    gen = gen or Ranlux(seed, luxury)
    nw = sum(dims[l] * dims[l + 1] for l in range(len(dims) - 1)) + dims[-1]
    nb = sum(dims)
    out = np.zeros((n_species, nw + nb))
    for sp in range(n_species):
        o = 0
        for l in range(len(dims) - 1):
            w = xavier_weights(gen, dims[l], dims[l + 1])
            out[sp, o:o + dims[l] * dims[l + 1]] = w.T.reshape(-1)  # column-major ww(:, o)
            o += dims[l] * dims[l + 1]
        xavier_weights(gen, dims[-1], 1)  # drawn, then zeroed (network.F90:120)
    return out, gen
