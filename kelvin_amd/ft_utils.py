"""Fermi-Dirac helpers for System providers (host side, O(n)).

Stand-ins for the cqcpy ``ft_utils`` functions the reference's systems call
(kelvin/ueg_system.py:36-45,126-131,400-419).  The hot path forms its own
FP64 Fermi weights on the device (libftcc ``ftcc_set_system``).
"""
import numpy


def ff(beta, en, mu):
    """Occupation 1/(1+exp(beta(e-mu))), evaluated with exp of -|x|."""
    x = beta * (numpy.asarray(en, dtype=numpy.float64) - mu)
    ex = numpy.exp(-numpy.abs(x))
    return numpy.where(x > 0, ex / (1.0 + ex), 1.0 / (1.0 + ex))


def ffv(beta, en, mu):
    """1 - ff, evaluated without cancellation."""
    x = beta * (numpy.asarray(en, dtype=numpy.float64) - mu)
    ex = numpy.exp(-numpy.abs(x))
    return numpy.where(x > 0, 1.0 / (1.0 + ex), ex / (1.0 + ex))


def GP0(beta, en, mu):
    """-(1/beta) ln(1 + exp(-beta(e-mu))) per orbital."""
    x = beta * (numpy.asarray(en, dtype=numpy.float64) - mu)
    return -(numpy.maximum(-x, 0.0) + numpy.log1p(numpy.exp(-numpy.abs(x)))) / beta


def dGP0(beta, en, mu):
    """d GP0 / d beta per orbital."""
    e = numpy.asarray(en, dtype=numpy.float64)
    return ((e - mu) * ff(beta, e, mu) - GP0(beta, e, mu)) / beta


def HtoK(T):
    """Hartree to Kelvin (the factor behind the temperatures the reference
    logs, e.g. examples/ueg_ft_ccsd_ESN.out:8)."""
    return T * 315775.12891
