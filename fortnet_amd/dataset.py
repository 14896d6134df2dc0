"""Reading structures from the arrays of an fnetdata file (``fnetdata.F90:994-1216``).

The arrays are the ones stored by the dataset conversion (one entry per datapoint, atoms
concatenated).  Fractional coordinates become Cartesian with ``matmul(latVecs, frac)``
(``fnetdata.F90:1147-1148``), accumulated in the same order.
"""
import numpy as np

from .core import Geometry


def frac_to_cart(frac, lat):
    frac = np.asarray(frac, np.float64)
    out = np.zeros_like(frac)
    for r in range(3):
        acc = np.zeros(frac.shape[0])
        for c in range(3):
            acc = acc + lat[c, r] * frac[:, c]
        out[:, r] = acc
    return out


def geometries(ds):
    """List of :class:`Geometry` of a converted dataset (species = 0-based network index)."""
    out = []
    off = ds["offsets"]
    for s in range(int(ds["ndatapoints"])):
        a, b = int(off[s]), int(off[s + 1])
        xyz = ds["coords"][a:b]
        lat = ds["basis"][s] if ds["periodic"][s] else None
        if ds["periodic"][s] and ds["fractional"][s]:
            xyz = frac_to_cart(xyz, ds["basis"][s])
        out.append(Geometry(xyz, ds["atnum"][a:b], ds["globalsp"][a:b] - 1, lat, ds["ext"][a:b]))
    return out
