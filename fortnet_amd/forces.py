"""Finite-difference forces (the reference's ``Forces = Numerical`` analysis).

``forceAnalysis_finiteDiff`` (``forces.F90:182-313``) displaces every Cartesian coordinate of
every atom by -delta and +delta (``TGeometriesForFiniteDiff_init``, ``forces.F90:85-117``; order
atom, coordinate, minus before plus), predicts the summed network outputs of all 6 N copies and
forms F[3 t + c, k] = (E_t(-delta) - E_t(+delta)) / (2 delta).  Here the copies of all requested
structures go to the device as batches of up to ``max_atoms`` atoms and their energies are
reduced there in FP64; the differences are taken on the host.
"""
import numpy as np

from ._lib import FNET_FP64
from .core import Geometry


def _displaced(g, delta):
    out = []
    for k in range(g.n):
        for c in range(3):
            for sgn in (-1.0, 1.0):
                xyz = g.coords.copy()
                xyz[k, c] = xyz[k, c] + sgn * delta
                out.append(Geometry(xyz, g.atnum, g.species,
                                    g.latvecs if g.periodic else None, g.ext))
    return out


def finite_difference_forces(ctx, geos, delta=0.01, precision=FNET_FP64, max_atoms=1 << 21):
    """Forces [n_k, 3 n_out] of every structure by central differences of the predicted
    energies (the model uploaded to ``ctx``); delta in Bohr (the reference's default 0.01)."""
    nout = ctx.model.n_out
    result = []
    pending = []  # (structure index, list of displaced geometries)
    energies = {}

    def flush():
        if not pending:
            return
        batch_geos = [d for _, ds in pending for d in ds]
        b = ctx.batch(batch_geos, precision)
        Es, _ = b.energy()
        b.close()
        o = 0
        for s, ds in pending:
            energies[s] = Es[o:o + len(ds)]
            o += len(ds)
        pending.clear()

    atoms = 0
    for s, g in enumerate(geos):
        ds = _displaced(g, delta)
        if atoms + len(ds) * g.n > max_atoms:
            flush()
            atoms = 0
        pending.append((s, ds))
        atoms += len(ds) * g.n
    flush()
    for s, g in enumerate(geos):
        E = energies[s].reshape(g.n, 3, 2, nout)
        F = (E[:, :, 0, :] - E[:, :, 1, :]) / (2.0 * delta)  # [atom, coordinate, output]
        result.append(np.ascontiguousarray(F.transpose(0, 2, 1).reshape(g.n, 3 * nout)))
    return result
