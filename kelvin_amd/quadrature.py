"""Imaginary-time grids and exp(-tau D) integrators.

Grid construction (ti, g, G for ng points, ng <= a few hundred) is host
bookkeeping with the same rules and names as kelvin/quadrature.py:17-256.
The integrators ``int_tbar1`` / ``int_tbar2`` (kelvin/quadrature.py:292-317)
run on the GPU through libftcc (``ftcc_integrate``).
"""
import numpy
import torch

from .engine import FTEngine


def get_G(ng, delta):
    """Simpson weights for int_0^{t_y} on a uniform grid
    (kelvin/quadrature.py:32-41): first panel trapezoid for odd rows."""
    G = numpy.zeros((ng, ng))
    if ng > 1:
        G[1, 0] = G[1, 1] = 0.5 * delta
    for y in range(2, ng):
        G[y] = G[y - 2]
        G[y, y - 2] += delta / 3.0
        G[y, y - 1] += 4.0 * delta / 3.0
        G[y, y] += delta / 3.0
    return G


def get_gint(ng, delta):
    """Simpson weights for int_0^beta (kelvin/quadrature.py:44-60)."""
    g = numpy.zeros(ng)
    if ng % 2 == 0:
        g[0] += 0.5 * delta
        g[1] += 0.5 * delta
        first = 3
    else:
        first = 2
    for y in range(first, ng, 2):
        g[y - 2] += delta / 3.0
        g[y - 1] += 4.0 * delta / 3.0
        g[y] += delta / 3.0
    return g


def get_G_midpoint(ng, delta):
    """kelvin/quadrature.py:17-23: lower-triangular (diagonal included)."""
    return numpy.tril(numpy.full((ng, ng), delta))


def get_g_midpoint(ng, delta):
    return numpy.full(ng, delta)


def get_gL(ng, delta):
    g = numpy.full(ng, delta)
    g[-1] = 0.0
    return g


def get_GL(ng, delta):
    """kelvin/quadrature.py:70-75: G[i,j] = delta for j < i-1."""
    return numpy.tril(numpy.full((ng, ng), delta), -2)


def _uniform(ng, beta):
    delta = beta / (ng - 1.0)
    return numpy.array([float(i) * delta for i in range(ng)]), delta


def simpsons(ng, beta):
    ti, delta = _uniform(ng, beta)
    return ti, get_gint(ng, delta), get_G(ng, delta)


def d_simpsons(ng, beta):
    delta = beta / (ng - 1.0)
    return get_gint(ng, delta / beta), get_G(ng, delta / beta)


def midpoint(ng, beta):
    delta = beta / ng
    ti = numpy.array([float(i) * delta + delta / 2 for i in range(ng)])
    return ti, get_g_midpoint(ng, delta), get_G_midpoint(ng, delta)


def left(ng, beta):
    ti, delta = _uniform(ng, beta)
    return ti, get_gL(ng, delta), get_GL(ng, delta)


def d_left(ng, beta):
    delta = beta / (ng - 1.0)
    return get_gL(ng, delta / beta), get_GL(ng, delta / beta)


def _mapped_rule(ng, quad):
    """(delta, s_i) of the mapped Simpson rules (kelvin/quadrature.py:119-212)."""
    if quad == 'ln':
        delta = (numpy.exp(1) - 1.0) / (ng - 1.0)
        return delta, numpy.asarray([(float(i) * delta + 1.0) for i in range(ng)])
    if quad == 'sin':
        delta = numpy.pi / (ng - 1.0)
        return delta, numpy.asarray([(float(i) * delta - numpy.pi / 2)
                                     for i in range(ng)])
    if quad == 'exp':
        delta = numpy.log(2.0) / (ng - 1.0)
    else:
        delta = 1.0 / (ng - 1.0)
    return delta, numpy.asarray([float(i) * delta for i in range(ng)])


def _power(quad):
    """The power n of the polynomial map t = beta (s^n + s) / 2: a rule
    name, or n itself (simpsons_p)."""
    if isinstance(quad, str):
        return {'quad': 2, 'cub': 3, 'quar': 4}[quad]
    return int(quad)


def _weigh(quad, x, si, scale):
    """x * scale * dt/ds in the reference's operation order."""
    if quad == 'ln':
        return scale * x / si
    if quad == 'sin':
        return scale * x * numpy.cos(si) / 2.0
    if quad == 'exp':
        return scale * x * numpy.exp(si)
    p = _power(quad)
    return scale * x * (numpy.power(si, p - 1) * float(p) + 1.0) / 2.0


def _times(quad, si, beta):
    if quad == 'ln':
        return beta * numpy.log(si)
    if quad == 'sin':
        return beta * (numpy.sin(si) + 1.0) / 2.0
    if quad == 'exp':
        return beta * (numpy.exp(si) - 1.0)
    p = _power(quad)
    return beta * (numpy.power(si, p) + si) / 2.0


def _mapped(ng, beta, quad, with_t):
    delta, si = _mapped_rule(ng, quad)
    g = _weigh(quad, get_gint(ng, delta), si, beta)
    G = get_G(ng, delta)
    for i in range(ng):
        G[i] = _weigh(quad, G[i], si, beta)
    if with_t:
        return _times(quad, si, beta), g, G
    return g, G


MAPPED = ('ln', 'sin', 'exp', 'quad', 'cub', 'quar')


# The named mapped rules of kelvin/quadrature.py:118-212: Simpson's rule in
# s, mapped to t by t = beta ln s (ln), beta (sin s + 1)/2 (sin),
# beta (e^s - 1) (exp) and beta (s^n + s)/2 (p); (ti, g, G), and the beta
# derivatives (g, G) of the d_ forms.
def simpsons_ln(ng, beta):
    return _mapped(ng, beta, 'ln', True)


def d_simpsons_ln(ng, beta):
    return _mapped(ng, 1.0, 'ln', False)


def simpsons_sin(ng, beta):
    return _mapped(ng, beta, 'sin', True)


def d_simpsons_sin(ng, beta):
    return _mapped(ng, 1.0, 'sin', False)


def simpsons_exp(ng, beta):
    return _mapped(ng, beta, 'exp', True)


def d_simpsons_exp(ng, beta):
    return _mapped(ng, 1.0, 'exp', False)


def simpsons_p(ng, beta, n=2):
    return _mapped(ng, beta, n, True)


def d_simpsons_p(ng, beta, n=2):
    return _mapped(ng, 1.0, n, False)


def ft_quad(ng, beta, quad):
    """Return (ti, g, G) (kelvin/quadrature.py:215-235)."""
    if quad == 'lin':
        return simpsons(ng, beta)
    if quad in MAPPED:
        return _mapped(ng, beta, quad, True)
    if quad == 'mid':
        return midpoint(ng, beta)
    if quad == 'L':
        return left(ng, beta)
    raise Exception("Unrecognized quadrature rule: {}".format(quad))


def d_ft_quad(ng, beta, quad):
    """Return d(g, G)/dbeta (kelvin/quadrature.py:238-256)."""
    if quad == 'lin':
        return d_simpsons(ng, beta)
    if quad in MAPPED:
        return _mapped(ng, 1.0, quad, False)
    if quad == 'L':
        return d_left(ng, beta)
    raise Exception("Unrecognized quadrature rule: {}".format(quad))


# ---------------------------------------------------------------------------
# GPU integrators
# ---------------------------------------------------------------------------
def _int_tbar(ng, tbar, ti, D1, D2, G, which):
    t = tbar if isinstance(tbar, torch.Tensor) else torch.as_tensor(
        numpy.ascontiguousarray(tbar, dtype=numpy.float64))
    if which == 1:
        nv, no = t.shape[1:]
    else:
        nv, no = t.shape[1], t.shape[3]
    eng = FTEngine(no, nv, ng)
    out = eng.to_device(t).clone()
    eng.set_grid(ti, G)
    if which == 1:
        d1 = D1
        d2 = torch.zeros((nv, nv, no, no), dtype=torch.float64)
        eng.set_denominators(d1, d2)
        eng.integrate(T1=out)
    else:
        d1 = torch.zeros((nv, no), dtype=torch.float64)
        eng.set_denominators(d1, D2)
        eng.integrate(T2=out)
    eng.close()
    return out


def int_tbar1(ng, t1bar, ti, D1, G):
    """T1(tau_y) = sum_x G[y,x] exp(-(tau_y-tau_x) D1) T1bar(tau_x) on the
    GPU; returns a new device tensor (kelvin/quadrature.py:292-303)."""
    return _int_tbar(ng, t1bar, ti, D1, None, G, 1)


def int_tbar2(ng, t2bar, ti, D2, G):
    """Doubles counterpart of int_tbar1 (kelvin/quadrature.py:306-317)."""
    return _int_tbar(ng, t2bar, ti, None, D2, G, 2)
