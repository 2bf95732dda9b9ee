"""Moment bases, exact moment matrices and Gram-Schmidt orthogonalisation.

Restates the parts of lbmpy's ``src/lbmpy/moments.py`` and
``src/lbmpy/methods/default_moment_sets.py`` the sweep depends on:

* ``moment_sort_key`` (``moments.py:175-179``): (order, number of symbols,
  length of the printed form, printed form).  The printed form is SymPy's
  ``str`` of a monomial, reproduced by :func:`monomial_str`.
* ``get_order`` (``moments.py:263-266``): total degree of the *lex-leading*
  monomial (SymPy ``LM`` over the symbols present), not the maximum degree.
* ``is_shear_moment`` / ``is_bulk_moment`` (``moments.py:369-405``).
* ``moment_matrix`` (``moments.py:417-440``), ``gram_schmidt`` (``:469-509``),
  ``get_default_moment_set_for_stencil`` (``:512-548``).
* ``cascaded_moment_sets_literature`` (``default_moment_sets.py:9-145``).

All arithmetic is exact (``fractions.Fraction``).
"""
from fractions import Fraction
from functools import reduce
from itertools import product
from math import gcd

from ._poly import Poly, MOMENT_VARS, moment_poly


# ------------------------------------------------------------------ printing & ordering

def monomial_str(e):
    """SymPy's ``str`` of the monomial x^a y^b z^c (factors in symbol order)."""
    parts = []
    for v, k in zip(MOMENT_VARS, e):
        if k == 1:
            parts.append(v)
        elif k > 1:
            parts.append(f"{v}**{k}")
    return "*".join(parts) if parts else "1"


def _as_exps(m, dim=3):
    e = tuple(m) + (0,) * (3 - len(m))
    return e


def monomial_sort_key(e):
    """``moment_sort_key`` for a single monomial (lbmpy ``moments.py:175-179``)."""
    e = _as_exps(e)
    s = monomial_str(e)
    return sum(e), sum(1 for k in e if k), len(s), s


exponent_tuple_sort_key = monomial_sort_key


def get_order(p):
    """Order of a moment polynomial as lbmpy computes it: degree of the lex-leading monomial."""
    if isinstance(p, tuple):
        return sum(p)
    if not p.terms:
        return 0
    present = [i for i in range(3) if any(e[i] for e in p.terms)]
    if not present:
        return 0
    lead = max(p.terms, key=lambda e: tuple(e[i] for i in present))
    return sum(lead)


def is_even(p):
    if isinstance(p, tuple):
        return sum(p) % 2 == 0
    return all(sum(e) % 2 == 0 for e in p.terms)


def is_bulk_moment(p, dim):
    """x^2+y^2(+z^2) plus lower orders (lbmpy ``moments.py:369-386``)."""
    if isinstance(p, tuple):
        p = moment_poly(p)
    quadratic = False
    found = [Fraction(0)] * dim
    for e, c in p.terms.items():
        if sum(e) == 2:
            quadratic = True
            for i in range(dim):
                if e[i] == 2:
                    found[i] += c
        elif sum(e) > 2:
            return False
    return quadratic and any(f != 0 for f in found) and len(set(found)) == 1


def is_shear_moment(p, dim):
    """Quadratic moments other than the bulk moment (lbmpy ``moments.py:389-405``)."""
    if is_bulk_moment(p, dim):
        return False
    if isinstance(p, tuple):
        p = moment_poly(p)
    quadratic = False
    for e in p.terms:
        if sum(e) == 2:
            quadratic = True
        elif sum(e) > 2:
            return False
    return quadratic


def sort_moments_into_groups_of_same_order(moments):
    """Groups keyed by order, in order of first appearance (lbmpy ``moments.py:182-188``)."""
    groups = {}
    for m in moments:
        groups.setdefault(get_order(m), []).append(m)
    return groups


# ------------------------------------------------------------------ exact linear algebra

def frac_matrix(rows):
    return [[Fraction(v) for v in r] for r in rows]


def mat_inv(a):
    """Exact inverse by Gauss-Jordan elimination."""
    n = len(a)
    m = [list(map(Fraction, row)) + [Fraction(int(i == j)) for j in range(n)] for i, row in enumerate(a)]
    for col in range(n):
        piv = next((r for r in range(col, n) if m[r][col] != 0), None)
        if piv is None:
            raise ValueError("singular matrix")
        m[col], m[piv] = m[piv], m[col]
        pv = m[col][col]
        m[col] = [v / pv for v in m[col]]
        for r in range(n):
            if r != col and m[r][col] != 0:
                f = m[r][col]
                m[r] = [a_ - f * b_ for a_, b_ in zip(m[r], m[col])]
    return [row[n:] for row in m]


def mat_mul(a, b):
    bt = list(zip(*b))
    return [[sum((x * y for x, y in zip(r, c)), Fraction(0)) for c in bt] for r in a]


def mat_vec(a, v):
    return [sum((x * y for x, y in zip(r, v)), Fraction(0)) for r in a]


# ------------------------------------------------------------------ moments on a stencil

def evaluate_moment(p, direction):
    """Polynomial moment evaluated at a lattice velocity."""
    d = tuple(direction) + (0,) * (3 - len(direction))
    total = Fraction(0)
    for e, c in p.terms.items():
        t = c
        for k, ci in zip(e, d):
            if k:
                t *= Fraction(ci) ** k
        total += t
    return total


def moment_matrix(moments, stencil):
    """Row i = moment i evaluated at every direction (lbmpy ``moments.py:417-440``)."""
    polys = [moment_poly(m) if isinstance(m, tuple) else m for m in moments]
    return [[evaluate_moment(p, d) for d in stencil] for p in polys]


def exponents_to_polys(exps):
    return [moment_poly(e) for e in exps]


def polynomial_monomials(p, dim):
    return {tuple(e[:dim]) for e in p.terms}


def common_denominator(p):
    dens = [c.denominator for c in p.terms.values()]
    return reduce(lambda a, b: a * b // gcd(a, b), dens, 1)


def gram_schmidt(moments, stencil, weights=None):
    """Weighted Gram-Schmidt on moment polynomials (lbmpy ``moments.py:469-509``)."""
    q = len(stencil)
    w = [Fraction(1)] * q if weights is None else [Fraction(x) for x in weights]
    polys = [moment_poly(m) if isinstance(m, tuple) else m for m in moments]
    cols = [[evaluate_moment(p, d) for d in stencil] for p in polys]
    ortho = []
    out = list(polys)
    for i in range(len(cols)):
        cur = list(cols[i])
        for j in range(i):
            prev = ortho[j]
            den = sum(p_ * p_ * w_ for p_, w_ in zip(prev, w))
            if den == 0:
                raise ValueError(f"Not an independent set of vectors given: vector {i} depends on previous ones")
            overlap = sum(c_ * p_ * w_ for c_, p_, w_ in zip(cur, prev, w)) / den
            cur = [c_ - overlap * p_ for c_, p_ in zip(cur, prev)]
            out[i] = out[i] - out[j] * overlap
        ortho.append(cur)
    return out


def all27_exponents():
    return list(product(range(3), repeat=3))


def get_default_moment_set_for_stencil(stencil):
    """Monomial moment set used to build equilibria (lbmpy ``moments.py:512-548``)."""
    D, Q = stencil.D, stencil.Q
    if D == 2 and Q == 9:
        exps = [(a, b, 0) for a, b in product(range(3), repeat=2)]
    elif D == 3 and Q == 27:
        exps = all27_exponents()
    elif D == 3 and Q == 19:
        removed = set()
        for base in [(1, 2, 2), (1, 1, 2), (2, 2, 2), (1, 1, 1)]:
            from itertools import permutations
            removed |= set(permutations(base))
        exps = [e for e in all27_exponents() if e not in removed]
    else:
        raise NotImplementedError("No default moment system available for this stencil")
    exps = sorted(exps, key=monomial_sort_key)
    return [moment_poly(e) for e in exps]


def cascaded_moment_sets_literature(stencil):
    """Polynomial groups for central-moment and cumulant methods
    (lbmpy ``default_moment_sets.py:9-145``)."""
    x, y, z = (Poly.var(MOMENT_VARS, v) for v in MOMENT_VARS)
    one = Poly.const(MOMENT_VARS, 1)
    D, Q = stencil.D, stencil.Q
    if D == 2 and Q == 9:
        return [[one], [x, y], [x * y, x ** 2 - y ** 2], [x ** 2 + y ** 2],
                [x ** 2 * y, x * y ** 2], [x ** 2 * y ** 2]]
    if D == 3 and Q == 19:
        return [[one], [x, y, z],
                [x * y, x * z, y * z, x ** 2 - y ** 2, x ** 2 - z ** 2],
                [x ** 2 + y ** 2 + z ** 2],
                [x * y ** 2 + x * z ** 2, x ** 2 * y + y * z ** 2, x ** 2 * z + y ** 2 * z],
                [x * y ** 2 - x * z ** 2, x ** 2 * y - y * z ** 2, x ** 2 * z - y ** 2 * z],
                [x ** 2 * y ** 2 - 2 * x ** 2 * z ** 2 + y ** 2 * z ** 2,
                 x ** 2 * y ** 2 + x ** 2 * z ** 2 - 2 * y ** 2 * z ** 2],
                [x ** 2 * y ** 2 + x ** 2 * z ** 2 + y ** 2 * z ** 2]]
    if D == 3 and Q == 27:
        return [[one], [x, y, z],
                [x * y, x * z, y * z, x ** 2 - y ** 2, x ** 2 - z ** 2],
                [x ** 2 + y ** 2 + z ** 2],
                [x * y ** 2 + x * z ** 2, x ** 2 * y + y * z ** 2, x ** 2 * z + y ** 2 * z],
                [x * y ** 2 - x * z ** 2, x ** 2 * y - y * z ** 2, x ** 2 * z - y ** 2 * z],
                [x * y * z],
                [x ** 2 * y ** 2 - 2 * x ** 2 * z ** 2 + y ** 2 * z ** 2,
                 x ** 2 * y ** 2 + x ** 2 * z ** 2 - 2 * y ** 2 * z ** 2],
                [x ** 2 * y ** 2 + x ** 2 * z ** 2 + y ** 2 * z ** 2],
                [x ** 2 * y * z, x * y ** 2 * z, x * y * z ** 2],
                [x ** 2 * y ** 2 * z, x ** 2 * y * z ** 2, x * y ** 2 * z ** 2],
                [x ** 2 * y ** 2 * z ** 2]]
    raise ValueError("No default set of cascaded moments is available for this stencil.")


def mrt_orthogonal_default_groups(stencil, weights):
    """Default MRT moment groups: weighted Gram-Schmidt of the default moments with the
    diagonal second-order moments replaced by (bulk, x^2, y^2-z^2), scaled to integer
    coefficients, grouped by order with shear and bulk split (lbmpy
    ``methods/creationfunctions.py:491-516``)."""
    x, y, z = (Poly.var(MOMENT_VARS, v) for v in MOMENT_VARS)
    moments = get_default_moment_set_for_stencil(stencil)
    if stencil.D == 2:
        diag = [x ** 2 + y ** 2, x ** 2]
    else:
        diag = [x ** 2 + y ** 2 + z ** 2, x ** 2, y ** 2 - z ** 2]
    for i, d in enumerate([x, y, z][:stencil.D]):
        sq = d ** 2
        if sq in moments:
            moments[moments.index(sq)] = diag[i]
    ortho = gram_schmidt(moments, stencil, weights)
    scaled = [p * common_denominator(p) for p in ortho]
    groups = list(sort_moments_into_groups_of_same_order(scaled).values())
    second = groups[2]
    bulk = [m for m in second if is_bulk_moment(m, stencil.D)]
    shear = [m for m in second if is_shear_moment(m, stencil.D)]
    assert len(bulk) + len(shear) == len(second)
    groups[2] = shear
    groups.insert(3, bulk)
    return groups


def monomial_to_polynomial_matrix(monomials, polynomials):
    """P with P[i][j] = coefficient of monomial j in polynomial i (lbmpy ``moments.py:583-609``)."""
    out = []
    for p in polynomials:
        row = [Fraction(0)] * len(monomials)
        for e, c in p.terms.items():
            row[monomials.index(tuple(e))] = c
        out.append(row)
    return out
