"""Deterministic synthetic workloads (structures, descriptor sets, networks).

All randomness comes from a counter-based SplitMix64 stream, so every machine regenerates the
same structures bit for bit.  Lengths are generated in Angstrom and stored in Bohr, the
reference's internal unit; the reference converts only ``RCut`` (``initprogram.F90:1354-1358``)
and uses eta as given, so eta in 1/Angstrom^2 is converted here: eta_Bohr = eta_AA * Bohr__AA^2.

Workloads (sizes are parameters so tests can use small instances):

* ``si_diamond``      -- n x n x m diamond Si supercells with Gaussian displacements;
* ``water_boxes``     -- rigid H2O (0.96 A, 104.5 deg) boxes, O sites by random sequential
                         addition with a minimum O-O distance;
* ``random_cells``    -- single-element cells, log-uniform atom counts, density +-5%, random
                         positions with a minimum pair distance;
* ``ternary_cells``   -- Si/C/O cells of the serving workload.
"""
import ctypes
import math
import os

import numpy as np

from .core import AA__BOHR, BOHR__AA, GFunctions, Geometry, Netstat, nparams

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64: uniform(n) returns the next n doubles in [0, 1)."""

    def __init__(self, seed):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.k = 0

    def uniform(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(self.k + 1, self.k + 1 + n, dtype=np.uint64)
            z = _mix(self.seed + idx * _GAMMA)
        self.k += n
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        """Box-Muller."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(1.0 - u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def eta_bohr(eta_aa):
    return np.asarray(eta_aa, np.float64) * BOHR__AA ** 2


# ---------------------------------------------------------------------------------------------
# structures

_DIAMOND = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


def si_diamond(nx, ny, nz, sigma_aa, seed, a_aa=5.43):
    """Diamond supercell, positions displaced by N(0, sigma) per component (Angstrom)."""
    cells = np.array([[i, j, k] for i in range(nx) for j in range(ny) for k in range(nz)], float)
    frac = (cells[:, None, :] + _DIAMOND[None, :, :]).reshape(-1, 3)
    xyz = frac * a_aa
    n = len(xyz)
    xyz = xyz + sigma_aa * Stream(seed).normal(3 * n).reshape(n, 3)
    lat = np.diag([nx * a_aa, ny * a_aa, nz * a_aa])
    return Geometry(xyz * AA__BOHR, np.full(n, 14), np.zeros(n, np.int32), lat * AA__BOHR)


def _min_image_d2(d, L):
    d = d - L * np.round(d / L)
    return (d * d).sum(-1)


_GEN = None


def _gen_lib():
    """libfnet_gen.so (fnet_gen.c, built by fortnet_amd.build), or None."""
    global _GEN
    if _GEN is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfnet_gen.so")
        _GEN = False
        if os.path.exists(path):
            lib = ctypes.CDLL(path)
            lib.fnet_rsa.restype = ctypes.c_int
            lib.fnet_rsa.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_long, ctypes.c_void_p]
            _GEN = lib
    return _GEN or None


def _rsa(st, n, L, dmin, max_tries=200000):
    """Random sequential addition of n points in a cubic box of side L with min distance.
    Candidates come 64 at a time; each is accepted if it keeps dmin to every point accepted
    before it.  The C helper (fnet_gen.c) gives bit-identical points to the numpy loop below
    (checked by the tests) in a fraction of the time."""
    lib = _gen_lib()
    if lib is not None and n > 0:
        pts = np.zeros((n, 3))
        k = ctypes.c_uint64(st.k)
        rc = lib.fnet_rsa(ctypes.c_uint64(int(st.seed)), ctypes.byref(k), n, float(L),
                          float(dmin), max_tries, pts.ctypes.data)
        if rc != 0:
            raise RuntimeError("random sequential addition did not converge")
        st.k = int(k.value)
        return pts
    return _rsa_numpy(st, n, L, dmin, max_tries)


def _rsa_numpy(st, n, L, dmin, max_tries=200000):
    """The numpy restatement of fnet_rsa: candidates of earlier batches checked for the whole
    batch at once, those accepted earlier in the same batch one by one."""
    pts = np.zeros((n, 3))
    k = 0
    tries = 0
    d2min = dmin * dmin
    while k < n:
        cand = st.uniform(3 * 64).reshape(64, 3) * L
        if k > 0:
            far = _min_image_d2(pts[None, :k, :] - cand[:, None, :], L).min(axis=1) >= d2min
        else:
            far = np.ones(64, bool)
        mutual = _min_image_d2(cand[None, :, :] - cand[:, None, :], L) >= d2min
        taken = []
        for c in range(64):
            tries += 1
            if tries > max_tries:
                raise RuntimeError("random sequential addition did not converge")
            if (k == 0 and not taken) or (far[c] and all(mutual[c, a] for a in taken)):
                pts[k] = cand[c]
                taken.append(c)
                k += 1
                if k == n:
                    break
    return pts


def _quat_rot(u):
    """Uniform random rotation from three uniforms (Shoemake)."""
    u1, u2, u3 = u
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    q = np.array([a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                  b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_box(seed, n_mol=64, L_aa=12.42, d_oh=0.96, angle_deg=104.5, dmin_oo=2.5):
    """Rigid water molecules in a cubic box; atoms ordered O, H, H per molecule.
    Species (network) indices: H = 0, O = 1 (dataset order of atomic numbers [1, 8])."""
    st = Stream(seed)
    O = _rsa(st, n_mol, L_aa, dmin_oo)
    half = math.radians(angle_deg) / 2
    h1 = d_oh * np.array([math.sin(half), 0.0, math.cos(half)])
    h2 = d_oh * np.array([-math.sin(half), 0.0, math.cos(half)])
    xyz, z = [], []
    for m in range(n_mol):
        R = _quat_rot(st.uniform(3))
        xyz += [O[m], O[m] + R @ h1, O[m] + R @ h2]
        z += [8, 1, 1]
    xyz = np.array(xyz)
    z = np.array(z)
    sp = np.where(z == 1, 0, 1)
    lat = np.eye(3) * L_aa
    return Geometry(xyz * AA__BOHR, z, sp, lat * AA__BOHR)


def random_cell(st, n, rho_gcc, dmin_aa, z, mass=28.0855):
    """Single-element cubic cell of n atoms at density rho (g/cm^3), min distance dmin (A).
    ``st`` is a Stream or a numpy Generator (tests)."""
    nd = rho_gcc / mass * 0.6022140857  # atoms per A^3
    L = (n / nd) ** (1.0 / 3.0)
    if isinstance(st, np.random.Generator):
        s2 = Stream(int(st.integers(0, 2 ** 62)))
    else:
        s2 = st
    xyz = _rsa(s2, n, L, dmin_aa)
    return Geometry(xyz * AA__BOHR, np.full(n, z), np.zeros(n, np.int32),
                    np.eye(3) * L * AA__BOHR)


def random_cells(n_struct, seed0, nmin=8, nmax=512, rho=2.33, dmin=2.0):
    out = []
    for s in range(n_struct):
        st = Stream(seed0 + s)
        n = int(math.floor(math.exp(math.log(nmin) + st.uniform(1)[0] *
                                    (math.log(nmax) - math.log(nmin)))))
        r = rho * (1.0 + (st.uniform(1)[0] * 0.1 - 0.05))
        out.append(random_cell(st, n, r, dmin, 14))
    return out


def ternary_cell(seed, nmin=16, nmax=4096, ndens=0.05, dmin=1.5):
    """Si/C/O cell of the serving workload; species indices in order [C, O, Si] = Z [6, 8, 14]."""
    st = Stream(seed)
    n = int(math.floor(math.exp(math.log(nmin) + st.uniform(1)[0] *
                                (math.log(nmax) - math.log(nmin)))))
    L = (n / ndens) ** (1.0 / 3.0)
    xyz = _rsa(st, n, L, dmin)
    zc = np.array([6, 8, 14])
    sp = np.minimum((st.uniform(n) * 3).astype(int), 2)
    return Geometry(xyz * AA__BOHR, zc[sp], sp.astype(np.int32), np.eye(3) * L * AA__BOHR)


# ---------------------------------------------------------------------------------------------
# descriptor sets and networks

def lj_energies(geos, eps_ev=0.1, sigma_aa=2.1, rc_aa=6.0):
    """Total Lennard-Jones energies (eV) of periodic structures, 4 eps ((s/r)^12 - (s/r)^6)
    summed over every pair of atoms and lattice images with r <= rc, each pair once: the
    training targets of the 64-atom Si task."""
    out = np.zeros(len(geos))
    rc = rc_aa * AA__BOHR
    for s, g in enumerate(geos):
        x = g.coords
        L = g.latvecs
        inv = np.linalg.inv(L)
        reach = [int(np.floor(rc * np.linalg.norm(inv[:, a]))) + 1 for a in range(3)]
        shifts = np.array([[i, j, k] for i in range(-reach[0], reach[0] + 1)
                           for j in range(-reach[1], reach[1] + 1)
                           for k in range(-reach[2], reach[2] + 1)], float) @ L
        e = 0.0
        for i in range(g.n):
            d = (x[None, :, :] + shifts[:, None, :]) - x[i][None, None, :]
            r = np.sqrt((d ** 2).sum(axis=2)).ravel() * BOHR__AA
            r = r[(r > 1e-8) & (r <= rc_aa)]
            sr6 = (sigma_aa / r) ** 6
            e += 0.5 * float(np.sum(4.0 * eps_ev * (sr6 * sr6 - sr6)))
        out[s] = e
    return out


def acsf_g2_g5(rc_aa, g2_eta_aa, g5_spec):
    """Unresolved G2 (rs = 0) + G5 table; g5_spec = list of (eta_aa, xi, lambda)."""
    rc = rc_aa * AA__BOHR
    n2, n5 = len(g2_eta_aa), len(g5_spec)
    return GFunctions([2] * n2 + [5] * n5, [rc] * (n2 + n5),
                      eta=np.concatenate([eta_bohr(g2_eta_aa), eta_bohr([e for e, _, _ in g5_spec])]),
                      lam=[0.0] * n2 + [l for _, _, l in g5_spec],
                      xi=[0.0] * n2 + [x for _, x, _ in g5_spec])


def random_params(dims, n_species, seed, scale=1.0):
    """Seeded Gaussian weights (std scale/sqrt(fan_in)), zero biases, dummy output layer zero."""
    st = Stream(seed)
    dims = [int(d) for d in dims]
    out = np.zeros((n_species, nparams(dims)))
    for sp in range(n_species):
        o = 0
        for l in range(len(dims) - 1):
            n = dims[l] * dims[l + 1]
            out[sp, o:o + n] = st.normal(n) * scale / math.sqrt(dims[l])
            o += n
    return out


def log_eta(lo, hi, n):
    return [lo * (hi / lo) ** (k / (n - 1)) for k in range(n)]


def config(name, scale=1.0):
    """(geometries, netstat without z-score, description) of a named workload.
    ``scale`` < 1 shrinks the number of structures / atoms for tests."""
    if name == "si64":
        n = max(1, int(round(100 * scale)))
        geos = [si_diamond(2, 2, 2, 0.1, 1 + s) for s in range(n)]
        g = acsf_g2_g5(6.0, log_eta(0.01, 2.0, 8),
                       [(0.005, xi, lam) for xi in (1, 2, 4, 16) for lam in (1.0, -1.0)])
        dims = [16, 20, 20, 1]
        ns = Netstat(g, [14], dims, 0, random_params(dims, 1, 0))
        desc = "%d x Si64 diamond (a=5.43 A, sigma=0.1 A), 8 G2 + 8 G5, Rc=6 A, 16-20-20-1" % n
    elif name == "water":
        n = max(1, int(round(500 * scale)))
        geos = [water_box(2000 + s) for s in range(n)]
        rc = 6.5 * AA__BOHR
        rad = GFunctions([2] * 10, [rc] * 10, eta=eta_bohr(log_eta(0.01, 2.0, 10)))
        ang = GFunctions([4] * 4, [rc] * 4, eta=eta_bohr([0.005] * 4), xi=[1, 1, 4, 4],
                         lam=[1.0, -1.0, 1.0, -1.0])
        g = GFunctions.concat([rad, ang]).species_resolved([1, 8])
        dims = [32, 25, 25, 1]
        ns = Netstat(g, [1, 8], dims, 0, random_params(dims, 2, 2))
        desc = ("%d x 64 H2O (192 atoms, 12.42 A box), species-resolved 2x10 G2 + 3x4 G4 = 32, "
                "Rc=6.5 A, H/O nets 32-25-25-1" % n)
    elif name == "si_cells":
        n = max(1, int(round(8192 * scale)))
        geos = random_cells(n, 3000)
        g = acsf_g2_g5(6.0, log_eta(0.01, 2.0, 8),
                       [(eta, xi, lam) for eta in (0.005, 0.02) for xi in (1, 2, 4, 8)
                        for lam in (1.0, -1.0)])
        dims = [24, 64, 64, 1]
        ns = Netstat(g, [14], dims, 0, random_params(dims, 1, 3))
        desc = "%d Si cells of 8-512 atoms, 8 G2 + 16 G5, Rc=6 A, 24-64-64-1" % n
    elif name == "si_bulk":
        nx = max(1, int(round(32 * scale)))
        nz = max(1, int(round(16 * scale)))
        geos = [si_diamond(nx, nx, nz, 0.05, 5)]
        g = acsf_g2_g5(6.0, log_eta(0.01, 2.0, 12),
                       [(eta, xi, lam) for eta in (0.005, 0.015) for xi in (1, 2, 4, 8, 16)
                        for lam in (1.0, -1.0)])
        dims = [32, 128, 128, 1]
        ns = Netstat(g, [14], dims, 0, random_params(dims, 1, 5))
        desc = "Si %dx%dx%d diamond (%d atoms), 12 G2 + 20 G5, Rc=6 A, 32-128-128-1" % (
            nx, nx, nz, geos[0].n)
    elif name == "ternary":
        n = max(1, int(round(64 * scale)))
        geos = [ternary_cell(4000 + s) for s in range(n)]
        rc = 6.0 * AA__BOHR
        rad = GFunctions([2] * 8, [rc] * 8, eta=eta_bohr(log_eta(0.01, 2.0, 8)))
        ang = GFunctions([5] * 4, [rc] * 4, eta=eta_bohr([0.005] * 4), xi=[1, 1, 4, 4],
                         lam=[1.0, -1.0, 1.0, -1.0])
        g = GFunctions.concat([rad, ang]).species_resolved([6, 8, 14])
        dims = [48, 64, 64, 1]
        ns = Netstat(g, [6, 8, 14], dims, 0, random_params(dims, 3, 4))
        desc = "%d ternary Si/C/O cells of 16-4096 atoms, 3x8 G2 + 6x4 G5 = 48, 48-64-64-1" % n
    else:
        raise ValueError(name)
    return geos, ns, desc


def with_zscore(ns, G):
    """Netstat with the z-score of ``TAcsf_getMeansAndVariances`` (acsf.F90:445-486) over G."""
    mean = G.mean(axis=0)
    std = np.sqrt(((G - mean) ** 2).mean(axis=0))
    return Netstat(ns.g, ns.species_z, ns.dims, ns.activation, ns.params, mean, std, ns.ext_in,
                   ns.n_global)
