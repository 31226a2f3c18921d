"""Line quadratures.

``gauss_hermite`` restates magritte/setup/quadrature.py:11-21: roots of the
physicists' Hermite polynomial H_n and weights 2^(n-1) n! / (n H_{n-1}(x))^2,
i.e. the Gauss-Hermite weights divided by sqrt(pi) so that they sum to one.

``gaussian_linspace`` is the quadrature the synthetic configurations use:
equally spaced roots over +-xmax Doppler widths with weights proportional to
exp(-x^2), normalised to one (the reference reads any Quadrature/{roots,weights},
src/Model/Lines/LineProducingSpecies/Quadrature/quadrature.cpp:19-40).
"""
from math import factorial

import numpy as np
from numpy.polynomial.hermite import hermroots, hermval


def gauss_hermite(n: int):
    roots = hermroots([0.0] * n + [1.0])
    h = hermval(roots, [0.0] * (n - 1) + [1.0])
    weights = 2 ** (n - 1) * factorial(n) / (n * h) ** 2
    return np.asarray(roots, dtype=np.float64), np.asarray(weights, dtype=np.float64)


def gaussian_linspace(n: int, xmax: float = 5.0):
    roots = np.linspace(-xmax, xmax, n)
    w = np.exp(-roots * roots)
    return roots.astype(np.float64), (w / w.sum()).astype(np.float64)
