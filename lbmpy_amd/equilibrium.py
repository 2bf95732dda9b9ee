"""Equilibrium distributions as exact polynomials in (rho, delta_rho, u).

Restates ``ContinuousHydrodynamicMaxwellian``
(lbmpy ``src/lbmpy/equilibrium/continuous_hydro_maxwellian.py:18-278``) and
``DiscreteHydrodynamicMaxwellian``
(``equilibrium/discrete_hydro_maxwellian.py:10-176`` with
``maxwellian_equilibrium.py:65-122``).  The continuous distribution's raw
moments are products of 1-D Gaussian moments G_n(u, cs^2) (the closed form of
``continuous_moment`` over the Maxwellian), truncated to ``order`` in u
exactly as ``_correct_order_and_cssq`` does.  Central moments are built by the
binomial expansion in raw moments (``_monomial_central_moment`` :206-221) and
cumulants are the Gaussian's (first order u, second order cs^2 on the
diagonal, zero above), rescaled by rho (:223-229).
"""
from fractions import Fraction
from math import comb

from ._poly import Poly
from .moments import moment_matrix, mat_inv, mat_vec, evaluate_moment

EQ_VARS = ("rho", "drho", "u0", "u1", "u2")
U_VARS = ("u0", "u1", "u2")


def _v(name):
    return Poly.var(EQ_VARS, name)


def _c(val):
    return Poly.const(EQ_VARS, val)


def gaussian_raw_1d(n, u, cs2):
    """E[xi^n] for xi ~ N(u, cs2): G_{n+1} = u G_n + n cs2 G_{n-1}."""
    g_prev, g = _c(0), _c(1)
    if n == 0:
        return g
    for k in range(n):
        g_prev, g = g, u * g + g_prev * (cs2 * k)
    return g


class ContinuousHydrodynamicMaxwellian:
    def __init__(self, dim=3, compressible=True, deviation_only=False, order=2, c_s_sq=Fraction(1, 3)):
        self.dim = dim
        self.compressible = compressible
        self.deviation_only = deviation_only
        self.order = order
        self.c_s_sq = Fraction(c_s_sq)
        self.continuous = True

    def _psi_moment(self, exps, density, u):
        val = density
        for d in range(self.dim):
            val = val * gaussian_raw_1d(exps[d] if d < len(exps) else 0, u[d], self.c_s_sq)
        return val

    def monomial_raw_moment(self, exps):
        u = [_v(U_VARS[d]) for d in range(self.dim)]
        zero = [_c(0)] * self.dim
        base_density = _v("rho") if self.compressible else _c(1)
        val = self._psi_moment(exps, base_density, u)
        if not self.compressible:
            val = val + self._psi_moment(exps, _v("drho"), zero)
        if self.deviation_only:
            val = val - self._psi_moment(exps, _c(1), zero)
        if self.order is not None:
            val = val.truncate(self.order, U_VARS)
        return val

    def moment(self, poly):
        """Equilibrium value of a (polynomial) raw moment."""
        total = _c(0)
        for e, c in poly.terms.items():
            total = total + self.monomial_raw_moment(e) * c
        return total

    def monomial_central_moment(self, exps):
        u = [_v(U_VARS[d]) for d in range(self.dim)] + [_c(0)] * (3 - self.dim)
        exps = tuple(exps) + (0,) * (3 - len(exps))
        val = _c(0)
        order = sum(exps)
        from itertools import product
        for rm in product(*(range(k + 1) for k in exps)):
            factor = _c((-1) ** (order - sum(rm)))
            for d in range(3):
                factor = factor * comb(exps[d], rm[d]) * (u[d] ** (exps[d] - rm[d]))
            val = val + factor * self.monomial_raw_moment(rm)
        if self.order is not None:
            val = val.truncate(self.order, U_VARS)
        return val

    def central_moment(self, poly):
        total = _c(0)
        for e, c in poly.terms.items():
            total = total + self.monomial_central_moment(e) * c
        return total

    def monomial_cumulant(self, exps):
        """rho * (Gaussian cumulant); the zeroth cumulant (rho log rho) is returned as None."""
        exps = tuple(exps) + (0,) * (3 - len(exps))
        if not self.compressible or self.deviation_only:
            raise ValueError("Cumulants can only be computed for the compressible, "
                             "non-deviation maxwellian equilibrium!")
        o = sum(exps)
        if o == 0:
            return None
        if o == 1:
            d = exps.index(1)
            return _v("rho") * _v(U_VARS[d])
        if o == 2 and 2 in exps:
            return _v("rho") * self.c_s_sq
        return _c(0)

    def cumulant(self, poly):
        total = _c(0)
        for e, c in poly.terms.items():
            v = self.monomial_cumulant(e)
            if v is None:
                return None
            total = total + v * c
        return total

    def background_moment(self, poly):
        """Moments of the rest state (rho=1, u=0) used for zero-centred storage."""
        total = Fraction(0)
        for e, c in poly.terms.items():
            v = c
            for d in range(self.dim):
                k = e[d]
                v *= 0 if k % 2 else (self.c_s_sq ** (k // 2)) * _double_factorial(k - 1)
            total += v
        return total


def _double_factorial(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def lattice_weights(stencil):
    """Hard-coded lattice weights (lbmpy ``maxwellian_equilibrium.py:24-61``)."""
    table = {9: {0: Fraction(4, 9), 1: Fraction(1, 9), 2: Fraction(1, 36)},
             7: {0: Fraction(0), 1: Fraction(1, 6)},
             15: {0: Fraction(2, 9), 1: Fraction(1, 9), 3: Fraction(1, 72)},
             19: {0: Fraction(1, 3), 1: Fraction(1, 18), 2: Fraction(1, 36)},
             27: {0: Fraction(8, 27), 1: Fraction(2, 27), 2: Fraction(1, 54), 3: Fraction(1, 216)}}
    return [table[stencil.Q][sum(abs(c) for c in d)] for d in stencil]


def get_weights(stencil, c_s_sq=Fraction(1, 3)):
    """lbmpy's name for the lattice weights (maxwellian_equilibrium.py:24-32); exact fractions.  They are the
    weights of the discrete equilibrium only for c_s^2 = 1/3, and a warning says so for any other value."""
    if c_s_sq != Fraction(1, 3) and str(c_s_sq) not in ("c_s**2", "1/3"):
        import warnings
        warnings.warn("Weights of discrete equilibrium are only valid if c_s^2 = 1/3")
    return lattice_weights(stencil)


class DiscreteHydrodynamicMaxwellian:
    """Textbook polynomial equilibrium w_i rho (1 + c.u/cs2 + (c.u)^2/(2 cs2^2) - u.u/(2 cs2))."""

    def __init__(self, stencil, compressible=True, deviation_only=False, order=2, c_s_sq=Fraction(1, 3)):
        self.stencil = stencil
        self.dim = stencil.D
        self.compressible = compressible
        self.deviation_only = deviation_only
        self.order = 4 if order is None else order
        self.c_s_sq = Fraction(c_s_sq)
        self.continuous = False
        w = lattice_weights(stencil)
        cs2 = self.c_s_sq
        u = [_v(U_VARS[d]) for d in range(self.dim)]
        uu = _c(0)
        for ua in u:
            uu = uu + ua * ua
        self.populations = []
        for wi, c in zip(w, stencil):
            cu = _c(0)
            for ca, ua in zip(c, u):
                cu = cu + ua * ca
            rho_in = _v("rho") if not compressible else _c(1)
            rho_out = _v("rho") if compressible else _c(1)
            fq = rho_in + cu * (1 / cs2)
            if self.order >= 2:
                fq = fq + cu * cu * (Fraction(1, 2) / cs2 ** 2) - uu * (Fraction(1, 2) / cs2)
            if self.order >= 3:
                fq = fq + cu * cu * cu * (Fraction(1, 6) / cs2 ** 3) - uu * cu * (Fraction(1, 2) / cs2 ** 2)
            val = fq * rho_out * wi
            if deviation_only:
                val = val - _c(wi)
            self.populations.append(val)

    def moment(self, poly):
        total = _c(0)
        for fi, d in zip(self.populations, self.stencil):
            total = total + fi * evaluate_moment(poly, d)
        return total

    def background_moment(self, poly):
        w = lattice_weights(self.stencil)
        return sum((wi * evaluate_moment(poly, d) for wi, d in zip(w, self.stencil)), Fraction(0))


def weights_by_moment_matching(stencil, moments, equilibrium):
    """Lattice weights as the discrete populations of the rest state (lbmpy
    ``momentbasedmethod.py:291-307``): M^-1 times the background moments."""
    M = moment_matrix(moments, stencil)
    bg = [equilibrium.background_moment(m) for m in moments]
    return mat_vec(mat_inv(M), bg)
