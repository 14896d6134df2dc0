"""The FP64 oracle against the reference's scratch-training fixtures (CPU).

For every steepest-descent, mse, unregularised case of ``training/*`` and ``input/*``: the
oracle's ACSF of the training set, the weighted z-score statistics (checked against the stored
ones), the RANLUX/Xavier initial weights, the oracle's sysTrain gradient summed over the
datapoints and normalised by the summed datapoint weights, and one steepest-descent step must
give the parameters the reference stored after its first iteration (ATOL 1e-10, RTOL 1e-9 of the
largest parameter).  The ``input/activation/*`` cases run the same chain with each of the ten
transfer functions, which pins the oracle's activations and their derivatives
(``transfer.F90:54-342``) on reference output.
"""
import numpy as np
import pytest

from conftest import case_names, load_case
from fortnet_amd.ranlux import initial_params
from oracle import oracle as O

ATOL, RTOL = 1e-10, 1e-9


def _cases():
    out = []
    for n in case_names():
        if not n.startswith(("input", "training")):
            continue
        c, _ = load_case(n)
        if (str(c["mode"]) == "train" and int(c["scratch"]) and str(c["optimizer"]) == "sd"
                and str(c["loss"]) == "mse" and not str(c["regu_type"])):
            out.append(n)
    return out


CASES = _cases()


def test_every_activation_is_covered():
    acts = set()
    for n in CASES:
        c, _ = load_case(n)
        acts.update(int(a) for a in c["after_activation"])
    assert acts == set(range(10))


@pytest.mark.parametrize("name", CASES)
def test_oracle_sd_step_from_scratch(name):
    c, ds = load_case(name)
    structs = O.dataset_structures(ds)
    off = ds["offsets"]
    w = ds["weights"].astype(np.float64)
    gf = None
    if "after_g_type" in c:
        g = lambda k: c["after_g_" + k]
        gf = O.GFunctions(g("type"), g("rcut"), g("eta"), g("rs"), g("kappa"), g("lam"),
                          g("xi"), g("atomid"), g("z1"), g("z2"))
    Gs = [O.acsf(st, gf) for st in structs] if gf is not None else None
    zm = zs = None
    if "after_zmean" in c:
        wa = np.repeat(w, np.diff(off))
        G = np.concatenate(Gs)
        zm = (wa[:, None] * G).sum(axis=0) / wa.sum()
        zs = np.sqrt((wa[:, None] * (G - zm) ** 2).sum(axis=0) / wa.sum())
        for got, want in ((zm, c["after_zmean"]), (zs, c["after_zstd"])):
            assert np.allclose(got, want, atol=ATOL, rtol=RTOL), np.abs(got - want).max()
    dims = c["after_topology"]
    nsp = len(c["after_net_atomicnumbers"])
    p0, _ = initial_params(dims, nsp, int(c["seed"]))
    nets = O.Nets(dims, c["after_activation"], p0)
    ng = int(c["after_nglobaltargets"])
    grad = np.zeros_like(p0)
    for s, st in enumerate(structs):
        parts = []
        if Gs is not None:
            parts.append(O.zscore(Gs[s], zm, zs))
        if "after_ext_indices" in c:
            parts.append(st.ext[:, c["after_ext_indices"] - 1])
        X = np.concatenate(parts, axis=1)
        a, b = int(off[s]), int(off[s + 1])
        sp = ds["globalsp"][a:b] - 1
        gt = ds["globaltargets"][s] if ng > 0 else None
        at = ds["atomictargets"][a:b] if int(dims[-1]) - ng > 0 else None
        nets.sys_train(X, sp, ng, gt, at, ds["atomicweights"][a:b], w[s], grad)
    grad /= w.sum()
    lr = float(c["learningrate"])
    maxdisp = float(c["maxdisp"]) if float(c["maxdisp"]) > 0 else 1e4
    new = p0.copy()
    for sp in range(nsp):
        step = -lr * grad[sp]
        m = np.abs(step).max()
        new[sp] = p0[sp] + (step if m <= maxdisp else (maxdisp / m) * step)
    want = c["after_params"]
    err = np.abs(new - want).max()
    assert np.allclose(new, want, atol=ATOL, rtol=RTOL), err
