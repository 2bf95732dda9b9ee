"""Exact multivariate polynomials over the rationals.

The method layer derives every table a sweep kernel needs (moment matrices,
equilibrium moments, lattice weights, force terms) in exact rational arithmetic
and only converts to floating point at the very end.  lbmpy does the same work
symbolically with SymPy (``src/lbmpy/moments.py``,
``src/lbmpy/equilibrium/continuous_hydro_maxwellian.py``); the quantities
needed here are all polynomials, so a small dict-of-monomials class is enough
and keeps the derivation fast and deterministic.

A polynomial is a mapping ``{exponent tuple: Fraction}`` over a fixed, ordered
tuple of variable names.
"""
from fractions import Fraction
from itertools import product


class Poly:
    __slots__ = ("vars", "terms")

    def __init__(self, variables, terms=None):
        self.vars = tuple(variables)
        self.terms = {}
        if terms:
            for e, c in terms.items():
                c = Fraction(c)
                if c != 0:
                    self.terms[tuple(e)] = c

    # ---- constructors -------------------------------------------------------
    @classmethod
    def const(cls, variables, c):
        return cls(variables, {(0,) * len(variables): c})

    @classmethod
    def var(cls, variables, name):
        idx = list(variables).index(name)
        e = [0] * len(variables)
        e[idx] = 1
        return cls(variables, {tuple(e): 1})

    @classmethod
    def monomial(cls, variables, exps, coeff=1):
        return cls(variables, {tuple(exps): coeff})

    # ---- arithmetic ---------------------------------------------------------
    def _coerce(self, other):
        if isinstance(other, Poly):
            if other.vars != self.vars:
                raise ValueError("polynomials over different variables")
            return other
        return Poly.const(self.vars, other)

    def __add__(self, other):
        other = self._coerce(other)
        t = dict(self.terms)
        for e, c in other.terms.items():
            t[e] = t.get(e, 0) + c
        return Poly(self.vars, t)

    __radd__ = __add__

    def __neg__(self):
        return Poly(self.vars, {e: -c for e, c in self.terms.items()})

    def __sub__(self, other):
        return self + (-self._coerce(other))

    def __rsub__(self, other):
        return self._coerce(other) - self

    def __mul__(self, other):
        if not isinstance(other, Poly):
            c = Fraction(other)
            return Poly(self.vars, {e: v * c for e, v in self.terms.items()})
        other = self._coerce(other)
        t = {}
        for e1, c1 in self.terms.items():
            for e2, c2 in other.terms.items():
                e = tuple(a + b for a, b in zip(e1, e2))
                t[e] = t.get(e, 0) + c1 * c2
        return Poly(self.vars, t)

    __rmul__ = __mul__

    def __pow__(self, n):
        r = Poly.const(self.vars, 1)
        for _ in range(n):
            r = r * self
        return r

    def __eq__(self, other):
        if not isinstance(other, Poly):
            other = Poly.const(self.vars, other)
        return self.vars == other.vars and self.terms == other.terms

    def __hash__(self):
        return hash((self.vars, frozenset(self.terms.items())))

    def is_zero(self):
        return not self.terms

    # ---- queries ------------------------------------------------------------
    def degree(self, var_subset=None):
        """Total degree, optionally counting only the given variables."""
        if not self.terms:
            return 0
        idx = range(len(self.vars)) if var_subset is None else [self.vars.index(v) for v in var_subset]
        return max(sum(e[i] for i in idx) for e in self.terms)

    def truncate(self, order, var_subset):
        """Drop every term whose degree in ``var_subset`` exceeds ``order``
        (lbmpy: ``remove_higher_order_terms(term, order, symbols=u)``)."""
        idx = [self.vars.index(v) for v in var_subset]
        return Poly(self.vars, {e: c for e, c in self.terms.items() if sum(e[i] for i in idx) <= order})

    def subs(self, mapping):
        """Substitute variables by numbers or polynomials (over the same variable tuple)."""
        result = Poly(self.vars)
        for e, c in self.terms.items():
            term = Poly.const(self.vars, c)
            for v, k in zip(self.vars, e):
                if k == 0:
                    continue
                if v in mapping:
                    val = mapping[v]
                    term = term * (val ** k if isinstance(val, Poly) else Fraction(val) ** k)
                else:
                    term = term * Poly.var(self.vars, v) ** k
            result = result + term
        return result

    def evaluate(self, values):
        """Evaluate at a point; ``values`` maps variable name -> number."""
        total = 0
        for e, c in self.terms.items():
            t = c
            for v, k in zip(self.vars, e):
                if k:
                    t = t * Fraction(values[v]) ** k if isinstance(values[v], (int, Fraction)) else t * values[v] ** k
            total = total + t
        return total

    def coeff(self, exps):
        return self.terms.get(tuple(exps), Fraction(0))

    def parity_pure(self):
        """True when every monomial has total degree of the same parity."""
        pars = {sum(e) % 2 for e in self.terms}
        return len(pars) <= 1

    def is_even(self):
        return all(sum(e) % 2 == 0 for e in self.terms)

    def monomials(self):
        return list(self.terms.keys())

    def __repr__(self):
        if not self.terms:
            return "0"
        parts = []
        for e, c in sorted(self.terms.items()):
            mono = "*".join(f"{v}**{k}" if k > 1 else v for v, k in zip(self.vars, e) if k)
            parts.append(f"{c}" + (f"*{mono}" if mono else ""))
        return " + ".join(parts)


MOMENT_VARS = ("x", "y", "z")


def moment_poly(exps):
    """Monomial moment x^a y^b z^c as a polynomial in the moment variables."""
    e = tuple(exps) + (0,) * (3 - len(exps))
    return Poly.monomial(MOMENT_VARS, e)


def all_exponents(max_component, dim):
    return list(product(range(max_component + 1), repeat=dim))
