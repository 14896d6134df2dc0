"""Pins the FP64 CPU oracle (CPU only).

The oracle is the checker of every GPU parity test, so it is itself checked here:

* against the reference's regression fixtures (predictions, validations, analytic forces,
  finite-difference forces, one steepest-descent step) at the reference's tolerance
  (ATOL 1e-10, RTOL 1e-9; testwithworkdir.py:24-25);
* against closed forms a plausible mistake would break (dimer G2, isolated trimers for G4/G5
  with the j == k diagonal of acsf.F90:1420-1432, cutoff values, diamond neighbour shells);
* against invariants (translation, rotation, permutation, supercell periodicity, sum of forces);
* analytic derivatives against central finite differences of the oracle's own values;
* the neighbour order against an independent brute-force enumeration.
"""
import math

import numpy as np
import pytest

from conftest import case_names, load_case
from oracle import oracle as O

ATOL, RTOL = 1e-10, 1e-9
PI = 3.14159265358979323846


def _close(a, b):
    return np.allclose(a, b, atol=ATOL, rtol=RTOL)


# prediction / validation runs of the reference's input-parsing tests with external atomic
# features: ACSF only ("empty" external list), external features only, and both mixed
EXT = [n for n in case_names("input__") if n.startswith(("input__extfeatures__prediction",
                                                           "input__extfeatures__validation",
                                                           "input__mixedfeatures__prediction",
                                                           "input__mixedfeatures__validation"))]
PRED = [n for n in case_names() if n.startswith(("analysis", "prediction", "validation"))] + EXT


@pytest.mark.parametrize("name", PRED)
def test_fixture_predictions_and_forces(name):
    c, ds = load_case(name)
    analytic = bool(c["out_tforces"]) and str(c["forces_kind"]) == "analytical"
    raw, glob, F = O.predict_case(c, ds, forces=analytic)
    assert _close(raw, c["out_raw"]), np.abs(raw - c["out_raw"]).max()
    if "out_glob" in c and c["out_glob"].size:
        ng = int(c["in_nglobaltargets"])
        assert _close(glob[:, :ng], c["out_glob"])
    if analytic:
        assert _close(F, c["out_forces"]), np.abs(F - c["out_forces"]).max()


@pytest.mark.parametrize("name", case_names("analysis__derivatives__numerical"))
def test_fixture_finite_difference_forces(name):
    c, ds = load_case(name)
    sts = O.dataset_structures(ds)
    off = ds["offsets"]
    F = np.concatenate([O.fd_forces(c, st, ds["globalsp"][off[s]:off[s + 1]] - 1,
                                    delta=float(c["delta"])) for s, st in enumerate(sts)])
    assert _close(F, c["out_forces"])


@pytest.mark.parametrize("name", ["analysis__derivatives__analytical__multiTargets__speciesResolved__raw",
                                  "analysis__derivatives__analytical__singleTarget__speciesUnresolved__atomID"])
def test_fixture_dense_forces(name):
    """The dense dG path (acsf.F90:870-939 + forces.F90:400-413) agrees with the fixtures too,
    in the exact image-sum mode and in the reference's scatter mode (no duplicate images here)."""
    c, ds = load_case(name)
    sts = O.dataset_structures(ds)
    off = ds["offsets"]
    for flags in (0, O.COMPAT_SCATTER | O.REFSHAPE):
        F = np.concatenate([O.dense_forces(c, st, ds["globalsp"][off[s]:off[s + 1]] - 1, flags)
                            for s, st in enumerate(sts)])
        assert _close(F, c["out_forces"])


def test_cos_regulariser_decides_force_parity():
    """The reference evaluates cos(theta) of the G4/G5 values as cos(acos(d / (r_j r_k + 1e-13)))
    (acsf.F90:1159-1176).  Dropping the 1e-13 (and taking j == k as exactly 1) moves each cosine
    by only ~1e-15, always in one direction; this fixture's z-scores have sigma down to 9e-8, so
    the shift reaches the forces at ~1e-9 and fails the reference's element-wise tolerance.  The
    planted mistake sits in the oracle (flag UNREG_COS); the device FP64 path keeps the
    regulariser for this reason."""
    name = "analysis__derivatives__analytical__multiTargets__speciesResolved__raw"
    c, ds = load_case(name)
    _, _, F = O.predict_case(c, ds, forces=True)
    assert _close(F, c["out_forces"])
    _, _, Fbad = O.predict_case(c, ds, flags=O.UNREG_COS, forces=True)
    assert not _close(Fbad, c["out_forces"])
    excess = np.abs(Fbad - c["out_forces"]) / (ATOL + RTOL * np.abs(c["out_forces"]))
    assert excess.max() > 2.0


SD = [n for n in case_names("restart") if n.endswith("steepestdesc") or "acsffeatures" in n]


@pytest.mark.parametrize("name", SD)
def test_fixture_sd_step(name):
    """Gradient (sysTrain/bprop/update normalisation) + steepest descent reproduce the netstat
    written after one iteration (steepdesc.F90:160-204 incl. the MaxDisplacement scaling)."""
    c, ds = load_case(name)
    if str(c["optimizer"]) != "sd":
        pytest.skip("not steepest descent")
    gf = O.GFunctions.from_case(c) if "in_g_type" in c else None
    nets = O.Nets.from_case(c)
    sts = O.dataset_structures(ds)
    off = ds["offsets"]
    ng = int(c["in_nglobaltargets"])
    grad = np.zeros_like(nets.params)
    for s, st in enumerate(sts):
        X = O.case_features(c, st, gf)
        nets.sys_train(X, ds["globalsp"][off[s]:off[s + 1]] - 1, ng,
                       ds["globaltargets"][s] if ng else None, ds["atomictargets"][off[s]:off[s + 1]],
                       ds["atomicweights"][off[s]:off[s + 1]], float(ds["weights"][s]), grad)
    grad /= ds["weights"].sum()
    lr = float(c["learningrate"])
    maxdisp = float(c["maxdisp"]) if float(c["maxdisp"]) > 0 else 1e4
    new = nets.params.copy()
    for sp in range(new.shape[0]):
        step = -lr * grad[sp]
        mx = np.abs(step).max()
        new[sp] += step if mx <= maxdisp else (maxdisp / mx) * step
    assert _close(new, c["after_params"]), np.abs(new - c["after_params"]).max()


# ---------------------------------------------------------------------------------------------
# closed forms

def _gf(types, rc, **kw):
    return O.GFunctions(types, [rc] * len(types), **kw)


def _iso(xyz, z=None):
    xyz = np.asarray(xyz, float)
    n = len(xyz)
    return O.Structure(xyz, None, 0, z if z is not None else np.full(n, 14))


def test_cutoff_and_g1_g2_g3_dimer():
    rc, r = 6.0, 2.3
    fc = 0.5 * (math.cos(PI * r / rc) + 1.0)
    st = _iso([[0, 0, 0], [r, 0, 0]])
    gf = _gf([1, 2, 3], rc, eta=[0, 0.7, 0], rs=[0, 1.1, 0], kappa=[0, 0, 1.3])
    G = O.acsf(st, gf)
    want = [fc, math.exp(-0.7 * (r - 1.1) ** 2) * fc, math.cos(1.3 * r) * fc]
    assert np.allclose(G[0], want, rtol=1e-14) and np.allclose(G[1], want, rtol=1e-14)
    # f_c(Rc) = 0, f_c(0) = 1: a dimer exactly at the cutoff contributes nothing
    G = O.acsf(_iso([[0, 0, 0], [rc, 0, 0]]), _gf([1], rc))
    assert abs(G[0, 0]) < 1e-15


def test_g5_g4_trimer_closed_form():
    """Equilateral trimer, side a: for the centre, two neighbours at r = a with cos = 1/2.
    Ordered pair sum incl. the diagonal (acsf.F90:1480-1488): 2 diagonal terms with cos ~ 1 and
    2 off-diagonal terms with cos = 1/2."""
    a, rc, eta, xi, lam = 2.0, 5.0, 0.3, 2.0, 1.0
    st = _iso([[0, 0, 0], [a, 0, 0], [a / 2, a * math.sqrt(3) / 2, 0]])
    fc = 0.5 * (math.cos(PI * a / rc) + 1)
    e = math.exp(-eta * 2 * a * a)
    cdiag = (a * a) / (a * a + 1e-13)
    g5 = 2 ** (1 - xi) * (2 * (1 + lam * cdiag) ** xi * e * fc * fc + 2 * (1 + lam * 0.5) ** xi * e * fc * fc)
    fjk = 0.5 * (math.cos(PI * a / rc) + 1)
    g4 = 2 ** (1 - xi) * (2 * (1 + lam * cdiag) ** xi * e * fc * fc * 1.0
                          + 2 * (1 + lam * 0.5) ** xi * e * math.exp(-eta * a * a) * fc * fc * fjk)
    G = O.acsf(st, _gf([5, 4], rc, eta=[eta, eta], xi=[xi, xi], lam=[lam, lam]))
    assert abs(G[0, 0] - g5) < 1e-13 * abs(g5)
    assert abs(G[0, 1] - g4) < 1e-13 * abs(g4)


def test_species_resolved_trimer():
    """C at the origin with one Si and one O neighbour: G5(Si,O) is a single ordered term
    (Z1 != Z2, no diagonal); G5(Si,Si) only the Si diagonal."""
    a, rc, eta, xi, lam = 2.0, 5.0, 0.1, 1.0, -1.0
    st = _iso([[0, 0, 0], [a, 0, 0], [0, a, 0]], np.array([6, 14, 8]))
    gf = _gf([5, 5], rc, eta=[eta, eta], xi=[xi, xi], lam=[lam, lam], z1=[14, 14], z2=[8, 14])
    G = O.acsf(st, gf)
    fc = 0.5 * (math.cos(PI * a / rc) + 1)
    e = math.exp(-2 * eta * a * a)
    cdiag = (a * a) / (a * a + 1e-13)
    assert abs(G[0, 0] - 2 ** (1 - xi) * (1 + lam * 0.0) ** xi * e * fc * fc) < 1e-14
    assert abs(G[0, 1] - 2 ** (1 - xi) * (1 + lam * cdiag) ** xi * e * fc * fc) < 1e-14


def test_diamond_neighbour_shells():
    """Perfect 2x2x2 Si diamond: shells 4 @ 2.35, 12 @ 3.84, 12 @ 4.50, 6 @ 5.43, 12 @ 5.92 A."""
    a = 5.43 / O.BOHR__AA
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    cells = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)])
    xyz = ((cells[:, None, :] + base[None]).reshape(-1, 3)) * a
    st = O.Structure(xyz, np.eye(3) * 2 * a, 1, np.full(64, 14))
    _, _, _, d2 = O.neighbours(st, 6.0 / O.BOHR__AA, 0)
    r = np.sort(np.sqrt(d2)) * O.BOHR__AA
    shells = np.round(r, 2)
    vals, counts = np.unique(shells, return_counts=True)
    assert list(vals) == [2.35, 3.84, 4.5, 5.43, 5.92]
    assert list(counts) == [4, 12, 12, 6, 12]


# ---------------------------------------------------------------------------------------------
# invariants and finite differences

def _random_case(seed, n=12, L=9.0, periodic=True, species=(14,)):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, L, size=(n, 3))
    z = np.asarray(species)[rng.integers(0, len(species), n)]
    lat = np.eye(3) * L if periodic else None
    return O.Structure(xyz, lat, int(periodic), z)


def _mixed_gf(rc, species=(6, 14)):
    types, z1, z2, eta, xi, lam, rs = [], [], [], [], [], [], []
    for z in species:
        for (t, e, r) in ((2, 0.3, 1.0), (2, 0.05, 0.0), (1, 0, 0), (3, 0, 0)):
            types.append(t); z1.append(z); z2.append(0); eta.append(e); rs.append(r)
            xi.append(0); lam.append(0)
    for a_ in range(len(species)):
        for b_ in range(a_, len(species)):
            for t in (4, 5):
                for (x, l) in ((1.0, 1.0), (2.0, -1.0), (3.5, 1.0)):
                    types.append(t); z1.append(species[a_]); z2.append(species[b_])
                    eta.append(0.02); xi.append(x); lam.append(l); rs.append(0)
    n = len(types)
    return O.GFunctions(types, [rc] * n, eta=eta, rs=rs, kappa=[0.7] * n, lam=lam, xi=xi,
                        z1=z1, z2=z2)


def test_rotation_translation_permutation_invariance():
    st = _random_case(3, 10, periodic=False, species=(6, 14))
    gf = _mixed_gf(5.0)
    G = O.acsf(st, gf)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    st2 = O.Structure(st.coords @ R.T + np.array([1.0, -2.0, 0.5]), None, 0, st.atnum)
    assert np.allclose(O.acsf(st2, gf), G, rtol=1e-10, atol=1e-12)
    perm = np.random.default_rng(1).permutation(st.n)
    st3 = O.Structure(st.coords[perm], None, 0, st.atnum[perm])
    assert np.allclose(O.acsf(st3, gf), G[perm], rtol=1e-10, atol=1e-12)


def test_supercell_periodicity():
    st = _random_case(4, 6, L=8.0, species=(6, 14))
    gf = _mixed_gf(6.5)
    G = O.acsf(st, gf)
    L = 8.0
    xyz = np.concatenate([st.coords + np.array([dx, 0, 0]) * L for dx in (0, 1)])
    big = O.Structure(xyz, np.diag([2 * L, L, L]), 1, np.concatenate([st.atnum] * 2))
    Gb = O.acsf(big, gf)
    assert np.allclose(Gb[:st.n], G, rtol=1e-9, atol=1e-11)
    assert np.allclose(Gb[st.n:], G, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("periodic", [False, True])
def test_derivatives_match_finite_differences(periodic):
    """Dense derivatives (gFuncGrad) against central differences of the oracle's values, in a
    cell short enough (periodic) that atoms see several images of each other and themselves."""
    st = _random_case(5, 5, L=5.5, periodic=periodic, species=(6, 14))
    gf = _mixed_gf(6.0)
    dG = O.acsf_prime(st, gf)  # [k][i][a][c]
    h = 1e-5
    for k in range(st.n):
        for c in range(3):
            xp, xm = st.coords.copy(), st.coords.copy()
            xp[k, c] += h
            xm[k, c] -= h
            Gp = O.acsf(O.Structure(xp, st.lat, st.periodic, st.atnum), gf)
            Gm = O.acsf(O.Structure(xm, st.lat, st.periodic, st.atnum), gf)
            fd = (Gp - Gm) / (2 * h)
            assert np.allclose(dG[k, :, :, c], fd, rtol=1e-6, atol=1e-7), (k, c)


def test_adjoint_equals_dense_and_sums_to_zero():
    st = _random_case(6, 9, L=7.0, species=(6, 14))
    gf = _mixed_gf(6.0)
    rng = np.random.default_rng(0)
    cG = rng.normal(size=(st.n, gf.n))
    F = O.acsf_adjoint(st, gf, cG)
    dG = O.acsf_prime(st, gf)
    Fd = -np.einsum("ia,kiac->kc", cG, dG)
    assert np.allclose(F, Fd, rtol=1e-11, atol=1e-12)
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * np.abs(F).max()


def test_compat_scatter_differs_only_with_duplicate_images():
    """Cell narrower than 2 Rc: the reference's last-write-wins scatter (acsf.F90:918) loses
    duplicate-image contributions; in a wide cell both modes agree."""
    gf = _mixed_gf(6.0, species=(14,))
    narrow = _random_case(8, 4, L=7.0, species=(14,))
    wide = _random_case(8, 4, L=14.0, species=(14,))
    assert np.allclose(O.acsf_prime(wide, gf), O.acsf_prime(wide, gf, O.COMPAT_SCATTER))
    assert not np.allclose(O.acsf_prime(narrow, gf), O.acsf_prime(narrow, gf, O.COMPAT_SCATTER))


def test_neighbour_order_against_brute_force():
    """Independent enumeration of (image, j) in the reference's order (dynneighlist.F90:233-320,
    latpointiter.F90:204-222): central cell j = i+1.., then images lexicographic, j = i.."""
    st = _random_case(9, 7, L=6.0, species=(14,))
    rc = 7.0
    lo, hi = O.image_ranges(st.lat, rc)
    for i in range(st.n):
        want = []
        for t in range(1, st.n):
            j = (i + t) % st.n
            d = st.coords[i] - st.coords[j]
            if (d * d).sum() <= rc * rc:
                want.append((j, -1))
        img = 0
        for p1 in range(lo[0], hi[0] + 1):
            for p2 in range(lo[1], hi[1] + 1):
                for p3 in range(lo[2], hi[2] + 1):
                    if p1 == p2 == p3 == 0:
                        continue
                    cell = st.lat.T @ np.array([p1, p2, p3], float)
                    for t in range(st.n):
                        j = (i + t) % st.n
                        d = st.coords[i] - (st.coords[j] + cell)
                        if (d * d).sum() <= rc * rc:
                            want.append((j, img))
                    img += 1
        j, im, _, _ = O.neighbours(st, rc, i)
        assert list(zip(j.tolist(), im.tolist())) == want
