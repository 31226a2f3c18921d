"""HEALPix RING-scheme pixel centres (Gorski et al. 2005, ApJ 622, 759).

The reference builds its ray set with ``healpy.pixelfunc.pix2vec(nside,
range(nrays))`` in RING ordering with weights 1/nrays
(magritte/setup/rays.py:54, magritte/setup/setup.py:224).  healpy is not
available here, so the pixel centres are computed from the published RING
formulas: polar caps z = +-(1 - i^2/(3 nside^2)), phi = (j - 1/2) pi / (2 i);
equatorial belt z = (2 nside - i) 2/(3 nside), phi = (j - s) pi / (2 nside)
with s = 1 when i + nside is odd and 1/2 otherwise.
"""
import math

import numpy as np


def npix(nside: int) -> int:
    return 12 * nside * nside


def pix2vec_ring(nside: int, ipix: int):
    """Unit vector of RING pixel ``ipix`` for resolution ``nside``."""
    if nside < 1:
        raise ValueError("nside must be >= 1")
    n_pix = npix(nside)
    if not 0 <= ipix < n_pix:
        raise ValueError("pixel index out of range")
    ncap = 2 * nside * (nside - 1)
    fact2 = 4.0 / n_pix
    fact1 = (2 * nside) * fact2
    if ipix < ncap:                                   # north polar cap
        iring = (1 + math.isqrt(1 + 2 * ipix)) >> 1
        iphi = (ipix + 1) - 2 * iring * (iring - 1)
        tmp = (iring * iring) * fact2
        z = 1.0 - tmp
        phi = (iphi - 0.5) * (0.5 * math.pi) / iring
        sth = math.sqrt(tmp * (2.0 - tmp))
    elif ipix < n_pix - ncap:                         # equatorial belt
        nl4 = 4 * nside
        tmp = ipix - ncap
        iring = tmp // nl4 + nside
        iphi = tmp % nl4 + 1
        fodd = 1.0 if ((iring + nside) & 1) else 0.5
        z = (2 * nside - iring) * fact1
        phi = (iphi - fodd) * math.pi * 0.75 * fact1
        sth = math.sqrt((1.0 - z) * (1.0 + z))
    else:                                             # south polar cap
        ip = n_pix - ipix
        iring = (1 + math.isqrt(2 * ip - 1)) >> 1
        iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1))
        tmp = (iring * iring) * fact2
        z = tmp - 1.0
        phi = (iphi - 0.5) * (0.5 * math.pi) / iring
        sth = math.sqrt(tmp * (2.0 - tmp))
    return (sth * math.cos(phi), sth * math.sin(phi), z)


def ring_rays(nside: int):
    """Directions [npix, 3] and equal weights 1/npix, as the reference sets up."""
    n = npix(nside)
    rays = np.array([pix2vec_ring(nside, p) for p in range(n)], dtype=np.float64)
    weights = np.full(n, 1.0 / n, dtype=np.float64)
    return rays, weights


def antipodes(rays: np.ndarray) -> np.ndarray:
    """Rays::setup_antipodal_rays (src/Model/Geometry/Rays/rays.cpp:90-103):
    the last r2 with |r1 + r2|^2 < 1e-9 is the antipode of r1 (-1 if none)."""
    n = rays.shape[0]
    out = np.full(n, -1, dtype=np.int64)
    for r1 in range(n):
        for r2 in range(n):
            s = rays[r1] + rays[r2]
            if (s[0] * s[0] + s[1] * s[1]) + s[2] * s[2] < 1.0e-9:
                out[r1] = r2
    return out
