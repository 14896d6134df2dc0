"""Synthetic workload generator (CPU only).

* The random sequential addition in C (libfnet_gen.so, fnet_gen.c) and its numpy restatement give
  bit-identical cells from the same stream (Si cells, ternary cells, water boxes).
* The Lennard-Jones targets of the 64-atom Si task: a dimer in a large box gives the pair
  formula; a perturbed periodic cell equals a brute-force sum over explicit images; a cell and
  its 2 x 1 x 1 supercell give energies in the ratio 2.
"""
import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import workloads as W


def test_rsa_c_matches_numpy():
    if W._gen_lib() is None:
        pytest.skip("libfnet_gen.so not built")
    for seed, n, L, dmin in [(3000, 37, 8.1, 2.0), (4001, 300, 18.0, 1.5), (7, 1, 5.0, 2.0),
                             (11, 100, 13.0, 2.2)]:
        a, b = W.Stream(seed), W.Stream(seed)
        pc = W._rsa(a, n, L, dmin)
        pn = W._rsa_numpy(b, n, L, dmin)
        assert np.array_equal(pc, pn)
        assert a.k == b.k  # the stream continues at the same place


def test_lj_dimer_and_periodic_cell():
    A = fa.AA__BOHR
    eps, sig = 0.1, 2.1
    for r in (1.9, 2.357, 3.3, 5.9):
        g = fa.Geometry(np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]]) * A, [14, 14], [0, 0],
                        np.eye(3) * 40.0 * A)
        e = W.lj_energies([g], eps, sig, 6.0)[0]
        assert abs(e - 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)) < 1e-12
    g = W.si_diamond(1, 1, 1, 0.1, 3)
    got = W.lj_energies([g])[0]
    x = g.coords * fa.BOHR__AA
    L = g.latvecs * fa.BOHR__AA
    want = 0.0
    for i in range(g.n):
        for j in range(g.n):
            for n in np.ndindex(5, 5, 5):
                if i == j and n == (2, 2, 2):
                    continue
                d = x[j] + (np.array(n) - 2) @ L - x[i]
                r = np.sqrt(d @ d)
                if r <= 6.0:
                    want += 0.5 * 4 * 0.1 * ((2.1 / r) ** 12 - (2.1 / r) ** 6)
    assert abs(got - want) < 1e-9 * abs(want)
    c1 = W.si_diamond(2, 2, 2, 0.0, 1)
    c2 = W.si_diamond(4, 2, 2, 0.0, 1)
    e1, e2 = W.lj_energies([c1, c2])
    assert abs(e2 - 2 * e1) < 1e-9 * abs(e2)
