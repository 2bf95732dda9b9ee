"""Numeric method tables for the HIP sweep kernels.

Turns a method description (``methods.py``) plus bound relaxation-rate values
into the flat table a kernel family reads (layouts in ``csrc/collide_moment.h``
and ``csrc/collide_cumulant.h``).  All derivation is exact (Fractions) until the
relaxation rates enter; from there numpy float64 is used, and the table is
cast to the kernel's compute type at upload.
"""
from fractions import Fraction

import numpy as np

from ._poly import Poly, MOMENT_VARS
from .equilibrium import EQ_VARS, lattice_weights
from .moments import mat_inv, moment_matrix, cascaded_moment_sets_literature, get_order, is_shear_moment

FAMILY_MOMENT, FAMILY_CENTRAL, FAMILY_CUMULANT = 0, 1, 2
STENCIL_CODES = {"D2Q9": 0, "D3Q19": 1, "D3Q27": 2}


class UnsupportedMethod(NotImplementedError):
    pass


def bind_rate(r, bindings):
    """Numeric value of a relaxation rate (number or SymPy expression with free symbols)."""
    if hasattr(r, "free_symbols") and r.free_symbols:
        missing = [s.name for s in r.free_symbols if s.name not in bindings]
        if missing:
            raise ValueError(f"relaxation rate symbol(s) {missing} must be passed to the kernel call")
        return float(r.subs({s: bindings[s.name] for s in r.free_symbols}))
    return float(r)


def family_of(method):
    return {"moment": FAMILY_MOMENT, "central_moment": FAMILY_CENTRAL, "cumulant": FAMILY_CUMULANT}[method.kind]


# ------------------------------------------------------------------------------- moment family

def _storage_equilibrium(method):
    """Per-direction equilibrium polynomials in storage form (zero-centred or absolute).

    Raw-moment space with a non-delta equilibrium subtracts the background moments
    before the backward transform (rawmomenttransforms.py:345-413); that is the same
    as relaxing towards the delta equilibrium, so the storage form is
    f_eq_full - w (zero-centred) or f_eq_full (absolute)."""
    st = method.stencil
    eq = method.equilibrium_distribution
    M = moment_matrix(list(method.moments), st)
    Minv = mat_inv(M)
    meq = list(method.moment_equilibrium_values)
    w = method.weights
    zc = method.zero_centered_pdfs
    if zc and not eq.deviation_only:
        bg = [eq.background_moment(p) for p in method.moments]
        meq = [m - b for m, b in zip(meq, bg)]
    elif not zc and eq.deviation_only:
        raise UnsupportedMethod("delta equilibrium requires zero-centred storage")
    feq = []
    for i in range(st.Q):
        p = Poly(EQ_VARS)
        for k in range(st.Q):
            if Minv[i][k] != 0:
                p = p + meq[k] * Minv[i][k]
        feq.append(p)
    return feq, M, Minv


def _split_equilibrium(feq_i, compressible):
    """f_eq = A + B*drho + lam * (L.u + Qd.u^2 + Qo.uu), lam = rho (compressible) or 1."""
    # express rho as 1 + drho
    rho = Poly.const(EQ_VARS, 1) + Poly.var(EQ_VARS, "drho")
    p = feq_i.subs({"rho": rho})
    A = p.coeff((0, 0, 0, 0, 0))
    B = p.coeff((0, 1, 0, 0, 0))
    lin = [p.coeff((0, 0) + tuple(int(a == d) for a in range(3))) for d in range(3)]
    quad_d = [p.coeff((0, 0) + tuple(2 * int(a == d) for a in range(3))) for d in range(3)]
    pairs = [(0, 1), (0, 2), (1, 2)]
    quad_o = [p.coeff((0, 0) + tuple(int(a in pr) for a in range(3))) for pr in pairs]
    # verify the remaining terms are exactly (drho * the u-part) for compressible, none otherwise
    rebuilt = Poly.const(EQ_VARS, A) + Poly.var(EQ_VARS, "drho") * B
    upart = Poly(EQ_VARS)
    for d in range(3):
        e = [0, 0, 0, 0, 0]
        e[2 + d] = 1
        upart = upart + Poly.monomial(EQ_VARS, e, lin[d])
        e = [0, 0, 0, 0, 0]
        e[2 + d] = 2
        upart = upart + Poly.monomial(EQ_VARS, e, quad_d[d])
    for (a, b), c in zip(pairs, quad_o):
        e = [0, 0, 0, 0, 0]
        e[2 + a] = 1
        e[2 + b] = 1
        upart = upart + Poly.monomial(EQ_VARS, e, c)
    lam = rho if compressible else Poly.const(EQ_VARS, 1)
    if not (rebuilt + lam * upart - p).is_zero():
        raise UnsupportedMethod("equilibrium is not of the form A + B drho + lam * quadratic(u)")
    return A, B, lin, quad_d, quad_o


def moment_table(method, bindings, dtype=np.float64, inputs=False):
    """Flat table for the moment-family kernel (layout MomentTab in collide_moment.h).  With ``inputs`` the
    equilibrium takes density and velocity from input fields, so m_eq - m of a conserved moment is no longer a
    known constant: conserved moments whose rate differs from their parity's base rate become active moments
    instead of being folded into G0."""
    st = method.stencil
    Q, D = st.Q, st.D
    feq, M, Minv = _storage_equilibrium(method)
    comp = method.compressible
    A = np.zeros(Q)
    B = np.zeros(Q)
    L = np.zeros((Q, 3))
    QD = np.zeros((Q, 3))
    QO = np.zeros((Q, 3))
    for i in range(Q):
        a, b, lin, qd, qo = _split_equilibrium(feq[i], comp)
        c = st[i] + (0,) * (3 - D)
        for d in range(3):
            if lin[d] != 0 and c[d] == 0:
                raise UnsupportedMethod("equilibrium has a linear term the kernel skips")
        for k, (x, y) in enumerate([(0, 1), (0, 2), (1, 2)]):
            if qo[k] != 0 and (c[x] == 0 or c[y] == 0):
                raise UnsupportedMethod("equilibrium has a cross term the kernel skips")
        if D == 2 and (qo[1] != 0 or qo[2] != 0):
            raise UnsupportedMethod("2-D equilibrium with z terms")
        A[i], B[i] = float(a), float(b)
        L[i] = [float(v) for v in lin]
        QD[i] = [float(v) for v in qd]
        QO[i] = [float(v) for v in qo]

    rates = np.array([bind_rate(r, bindings) for r in method.relaxation_rates])
    polys = list(method.moments)
    if not all(p.parity_pure() for p in polys):
        raise UnsupportedMethod("moment polynomials of mixed parity")
    even = np.array([p.is_even() for p in polys])
    order = np.array([get_order(p) for p in polys])
    conserved = np.array([all(sum(e) <= 1 for e in p.terms) for p in polys])

    def base(mask):
        vals = rates[mask & ~conserved]
        if len(vals) == 0:
            vals = rates[mask]
        uniq, counts = np.unique(vals, return_counts=True)
        return float(uniq[np.argmax(counts)])

    r_even, r_odd = base(even), base(~even)
    Mf = np.array([[float(v) for v in r] for r in M])
    Minvf = np.array([[float(v) for v in r] for r in Minv])
    Dm = np.diag(rates)

    # force: affine coefficients of the force term in u, and the constant contributions of
    # conserved moments whose (m_eq - m) is known exactly
    G0 = np.zeros(Q)
    G = np.zeros((Q, 3))
    fm = method.force_model
    F = np.zeros(3)
    if fm is not None:
        F[:D] = fm.force
        t0, tu = fm.population_terms(method)
        t0 = np.array([float(v) for v in t0])
        tu = np.array([[float(v) for v in r] for r in tu])
        if type(fm).__name__ == "Guo":
            Acorr = Minvf @ (np.eye(Q) - 0.5 * Dm) @ Mf
            t0 = Acorr @ t0
            tu = Acorr @ tu
        G0 += t0
        G[:, :D] += tu

    def split(r_even, r_odd):
        """Moments relaxed individually (rate off their parity's base rate), and the constant force column
        of the conserved ones among them."""
        active, g0 = [], np.zeros(Q)
        for k in range(Q):
            rb = r_even if even[k] else r_odd
            dr = rates[k] - rb
            if dr == 0.0:
                continue
            if conserved[k] and not inputs:
                # m_eq - m for a conserved moment: 0 for the density part; F/2 for momentum when the
                # equilibrium velocity is shifted by F/(2 rho_ref)
                lin_coef = np.zeros(3)
                for e, cf in polys[k].terms.items():
                    if sum(e) == 1:
                        lin_coef[e.index(1)] = float(cf)
                dm = 0.0
                if fm is not None and fm.equilibrium_velocity_shift:
                    dm = float(lin_coef @ (0.5 * F))
                g0 += dr * Minvf[:, k] * dm
            else:
                active.append((k, dr))
        return active, g0

    active, g0 = split(r_even, r_odd)
    shear = _shear_projector(st, active, Mf, Minvf)
    if not shear[0]:
        # the same collision with the even base rate taken from the moments that are not shear moments: when the
        # shear moments are then the only active ones, the kernel applies their projector in closed form
        non_shear = even & ~conserved & ~np.array([is_shear_moment(p, D) for p in polys])
        if non_shear.any():
            uniq, counts = np.unique(rates[non_shear], return_counts=True)
            r_alt = float(uniq[np.argmax(counts)])
            act2, g02 = split(r_alt, r_odd)
            sh2 = _shear_projector(st, act2, Mf, Minvf)
            if sh2[0]:
                r_even, active, g0, shear = r_alt, act2, g02, sh2
    G0 += g0
    tab = np.zeros(16 * Q + 2 + 2 * Q * Q + 1 + Q)
    tab[16 * Q + 2 + 2 * Q * Q:] = shear
    tab[0:Q] = A
    tab[Q:2 * Q] = B
    tab[2 * Q:5 * Q] = L.reshape(-1)
    tab[5 * Q:8 * Q] = QD.reshape(-1)
    tab[8 * Q:11 * Q] = QO.reshape(-1)
    tab[11 * Q:12 * Q] = G0
    tab[12 * Q:15 * Q] = G.reshape(-1)
    tab[15 * Q] = r_even
    tab[15 * Q + 1] = r_odd
    for n, (k, dr) in enumerate(active):
        tab[15 * Q + 2 + n] = dr
        tab[16 * Q + 2 + n * Q: 16 * Q + 2 + (n + 1) * Q] = Mf[k]
        tab[16 * Q + 2 + Q * Q + n * Q: 16 * Q + 2 + Q * Q + (n + 1) * Q] = Minvf[:, k]
    return tab.astype(dtype), len(active)


def _shear_projector(st, active, Mf, Minvf):
    """The table's SHEAR block: [1, s (9/2) w_i ...] when the sum of the active moments' rank-one corrections,
    sum_k dr_k Minv[:, k] M[k], equals s times the w-weighted projector onto the traceless second-order (shear)
    moments, P_ij = (9/2) w_i (c_i c_i) : (c_j c_j - |c_j|^2 I / 3) -- as for an orthogonal MRT whose only
    rates off their parity's base rate are the shear moments' common rate (the default orthogonal D3Q19 / D3Q27
    MRT given its shear rate).  The kernel then applies s (9/2) w_i (c_i c_i) : Pi_dev(f_eq - f) in closed form.
    Zeros otherwise (3-D stencils only; checked to 1e-12 of the largest entry)."""
    Q, D = st.Q, st.D
    out = np.zeros(Q + 1)
    if D != 3 or not active:
        return out
    C = np.zeros((Q, Q))
    for k, dr in active:
        C += dr * np.outer(Minvf[:, k], Mf[k])
    w = np.array([float(x) for x in lattice_weights(st)])
    c = np.array([list(st[i]) for i in range(Q)], dtype=float)
    cc = np.einsum("ia,ib->iab", c, c)
    dev = cc - np.einsum("i,ab->iab", (c * c).sum(1) / 3.0, np.eye(3))
    P = 4.5 * w[:, None] * np.einsum("iab,jab->ij", cc, dev)
    scale = float((C * P).sum() / (P * P).sum())
    if np.abs(C - scale * P).max() > 1e-12 * max(1.0, np.abs(C).max()):
        return out
    out[0] = 1.0
    out[1:] = scale * 4.5 * w
    return out


def moment_shear_projector(tab, q):
    """Whether a moment-family table carries the closed-form shear projector (LBM355_FLAG_SHEAR_PROJECTOR)."""
    return tab.size == 16 * q + 2 + 2 * q * q + 1 + q and tab[16 * q + 2 + 2 * q * q] == 1.0


# ------------------------------------------------------------------------------- cumulant family

# ------------------------------------------------------------------------------- cumulant corrections

CORRECTION_RATE_NAMES = tuple(f"corr_rr_{i}" for i in range(10))
# which of the ten fourth-order rate symbols each polynomial of the default D3Q27 set carries, as
# _get_relaxation_info_dict assigns ten rates (shear, bulk, then one per remaining group);
# None: conserved (rate 0)
_CORRECTION_RATE_SLOTS = (None,) * 4 + (0,) * 5 + (1,) + (2,) * 3 + (3,) * 3 + (4,) + (5,) * 2 + (6,) + \
    (7,) * 3 + (8,) * 3 + (9,)


class CumulantCorrections:
    """The D3Q27 cumulant corrections an LBMConfig asks for (lbmpy ``creationfunctions.py:692-707``):
    the Galilean correction (``methods/cumulantbased/galilean_correction.py:21-91``) and the fourth-order
    correction (``methods/cumulantbased/fourth_order_correction.py:21-237``) with its limiter (1e6 for
    ``True``) and the shear / bulk rates its optimal parametrisation is computed from."""

    def __init__(self, galilean=False, fourth_order=False, limiter=1e6, shear=None, bulk=1):
        self.galilean = bool(galilean)
        self.fourth_order = bool(fourth_order)
        self.limiter = float(limiter)
        self.shear = shear
        self.bulk = bulk

    def free_symbols(self):
        out = []
        for r in (self.shear, self.bulk) if self.fourth_order else ():
            for s in getattr(r, "free_symbols", ()):
                if s.name not in out:
                    out.append(s.name)
        return out

    def __repr__(self):
        return (f"CumulantCorrections(galilean={self.galilean}, fourth_order={self.fourth_order}, "
                f"limiter={self.limiter}, shear={self.shear}, bulk={self.bulk})")


def cumulant_corrections(lbm_config):
    """The corrections of ``lbm_config`` or None.  The fourth-order correction takes the shear rate from
    ``relaxation_rates[0]`` and the bulk rate from ``relaxation_rates[1]`` (1 when only one is given,
    creationfunctions.py:699-701)."""
    gal = bool(lbm_config.galilean_correction)
    foc = lbm_config.fourth_order_correction
    if not gal and not foc:
        return None
    rr = list(lbm_config.relaxation_rates or [])
    if foc:
        limiter = 1e6 if foc is True else float(foc)
        return CumulantCorrections(gal, True, limiter, rr[0], rr[1] if len(rr) > 1 else 1)
    return CumulantCorrections(gal, False)


def fourth_order_parameters(w1, w2):
    """Optimal a, b (Geier et al. 2017, eq. 115-116) and unlimited omega_3, omega_4, omega_5 (eq. 112-114)
    for shear rate w1 and bulk rate w2, as lbmpy evaluates them (fourth_order_correction.py:70-94,
    127-139)."""
    d = (w1 - w2) * (w2 * (2.0 + 3.0 * w1) - 8.0 * w1)
    if d == 0.0:
        raise ValueError("the fourth-order correction needs distinct shear and bulk rates with "
                         "w2 (2 + 3 w1) != 8 w1")
    a = (4.0 * w1 * w1 + 2.0 * w1 * w2 * (w1 - 6.0) + w2 * w2 * (w1 * (10.0 - 3.0 * w1) - 4.0)) / d
    b = (4.0 * w1 * w2 * (9.0 * w1 - 16.0) - 4.0 * w1 * w1 - 2.0 * w2 * w2 * (2.0 + 9.0 * w1 * (w1 - 2.0))) / (3.0 * d)
    w3 = 8.0 * (w1 - 2.0) * (w2 * (3.0 * w1 - 1.0) - 5.0 * w1) / \
        (8.0 * (5.0 - 2.0 * w1) * w1 + w2 * (8.0 + w1 * (9.0 * w1 - 26.0)))
    w4 = 8.0 * (w1 - 2.0) * (w1 + w2 * (3.0 * w1 - 7.0)) / (w2 * (56.0 - 42.0 * w1 + 9.0 * w1 * w1) - 8.0 * w1)
    w5 = 24.0 * (w1 - 2.0) * (4.0 * w1 * w1 + w1 * w2 * (18.0 - 13.0 * w1) + w2 * w2 * (2.0 + w1 * (6.0 * w1 - 11.0))) / \
        (16.0 * w1 * w1 * (w1 - 6.0) - 2.0 * w1 * w2 * (216.0 + 5.0 * w1 * (9.0 * w1 - 46.0)) +
         w2 * w2 * (w1 * (3.0 * w1 - 10.0) * (15.0 * w1 - 28.0) - 48.0))
    return a, b, w3, w4, w5


# monomial groups of the general cumulant relaxation, in the order of CumulantTab's BLOCKS
_CUMULANT_BLOCKS = (
    ((2, 0, 0), (0, 2, 0), (0, 0, 2)),
    ((1, 1, 0), (1, 0, 1), (0, 1, 1)),
    ((1, 2, 0), (1, 0, 2), (2, 1, 0), (0, 1, 2), (2, 0, 1), (0, 2, 1)),
    ((1, 1, 1),),
    ((2, 2, 0), (2, 0, 2), (0, 2, 2)),
    ((2, 1, 1), (1, 2, 1), (1, 1, 2)),
    ((2, 2, 1), (2, 1, 2), (1, 2, 2)),
    ((2, 2, 2),),
)
_CONSERVED = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))


def cumulant_is_default_set(method):
    """True for lbmpy's default polynomial D3Q27 cumulant set (cascaded_moment_sets_literature), which the
    cumulant sweep relaxes with per-polynomial rates; other sets take the general kernel variant."""
    reference = [p for g in cascaded_moment_sets_literature(method.stencil) for p in g]
    return list(method.cumulants) == reference


def cumulant_blocks(method, rates):
    """The relaxation of the monomial cumulant deviations as linear maps per monomial group: with P the method's
    polynomial cumulants in monomial coordinates and D their rates, post = P^-1 (I - D) P pre (the equilibrium
    deviation of every cumulant is zero, cumulantbasedmethod.py:374-379).  Returns the blocks of
    _CUMULANT_BLOCKS flattened row-major; raises UnsupportedMethod when the set couples monomials of different
    groups, or relaxes a conserved (order 0 / 1) cumulant."""
    monos = [tuple(e) for e in _CANON]
    index = {e: i for i, e in enumerate(monos)}
    polys = list(method.cumulants)
    P = np.zeros((27, 27))
    for r, p in enumerate(polys):
        for e, cf in p.terms.items():
            P[r, index[tuple(e[:3]) + (0,) * (3 - len(e[:3]))]] += float(cf)
    if abs(np.linalg.det(P)) < 1e-12:
        raise UnsupportedMethod("the cumulant polynomials are not independent")
    M = np.linalg.solve(P, np.diag(1.0 - np.asarray(rates, dtype=np.float64)) @ P)
    for e in _CONSERVED:
        i = index[e]
        row, col = M[i].copy(), M[:, i].copy()
        row[i] -= 1.0
        col[i] -= 1.0
        if np.abs(row).max() > 1e-12 or np.abs(col).max() > 1e-12:
            raise UnsupportedMethod("the sweep keeps the conserved (order 0 and 1) cumulants unrelaxed")
    covered = np.zeros((27, 27), dtype=bool)
    for e in _CONSERVED:
        covered[index[e], index[e]] = True
    out = []
    for group in _CUMULANT_BLOCKS:
        ids = [index[e] for e in group]
        covered[np.ix_(ids, ids)] = True
        out.extend(M[np.ix_(ids, ids)].reshape(-1))
    if np.abs(M[~covered]).max(initial=0.0) > 1e-12:
        raise UnsupportedMethod("the cumulant set couples monomials the sweep relaxes in separate groups")
    return np.array(out)


def cumulant_table(method, bindings, dtype=np.float64, corrections=None):
    """Per-polynomial rates (for the default D3Q27 polynomial cumulant set, in its order), then the
    fourth-order correction's a, b, omega_3..5 and limiter, then the relaxation as linear maps on groups of
    monomial cumulants (cumulant_blocks; what the general kernel variant reads) -- layout CumulantTab in
    csrc/collide_cumulant.h."""
    st = method.stencil
    if st.name != "D3Q27":
        raise UnsupportedMethod("the cumulant sweep is implemented for D3Q27")
    default_set = cumulant_is_default_set(method)
    if not default_set and corrections is not None:
        raise UnsupportedMethod("the Galilean / fourth-order corrections are defined for the default polynomial "
                                "cumulant set")
    fm = method.force_model
    if fm is not None and type(fm).__name__ != "CentralMoment":
        raise UnsupportedMethod("cumulant sweep supports the CentralMoment force model")
    extra = np.zeros(6)
    named = {}
    if corrections is not None and corrections.fourth_order:
        w1 = bind_rate(corrections.shear, bindings)
        w2 = bind_rate(corrections.bulk, bindings)
        a, b, w3, w4, w5 = fourth_order_parameters(w1, w2)
        extra[:] = (a, b, w3, w4, w5, corrections.limiter)
        named = dict(zip(CORRECTION_RATE_NAMES, (w1, w2, w3, w4, w5, 1.0, 1.0, 1.0, 1.0, 1.0)))
        for slot, r in zip(_CORRECTION_RATE_SLOTS, method.relaxation_rates):
            want = 0.0 if slot is None else CORRECTION_RATE_NAMES[slot]
            got = getattr(r, "name", None) if slot is not None else float(r)
            if got != want:
                raise UnsupportedMethod("fourth-order correction: the method's rates are not the ten "
                                        "corr_rr symbols in the default polynomial groups")
    rates = np.array([named[r.name] if getattr(r, "name", None) in named else bind_rate(r, bindings)
                      for r in method.relaxation_rates])
    if rates[0] != 0.0:
        raise UnsupportedMethod("the zeroth cumulant must not be relaxed")
    if corrections is not None and corrections.galilean and rates[7] != rates[8]:
        raise ValueError("Cumulants (x^2 - y^2) and (x^2 - z^2) must have the same relaxation rate")
    blocks = cumulant_blocks(method, rates)
    if not default_set:
        rates = np.zeros(27)
    return np.concatenate([rates, extra, blocks]).astype(dtype), 0


# ------------------------------------------------------------------------------- central-moment family

_CANON = [(a, b, c) for a in range(3) for b in range(3) for c in range(3)]   # index 9a + 3b + c
_UMONO = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1),
          (0, 1, 1)]


def _gamma(e):
    g = Fraction(1)
    for k in e:
        g *= {0: Fraction(1), 1: Fraction(0), 2: Fraction(1, 3)}[k]
    return g


def central_table(method, bindings, dtype=np.float64):
    """Table for the central-moment kernel (layout CentralTab in csrc/collide_central.h)."""
    st = method.stencil
    D, Q = st.D, st.Q
    eq = method.equilibrium_distribution
    if eq.c_s_sq != Fraction(1, 3):
        raise UnsupportedMethod("c_s^2 must be 1/3")
    if not eq.continuous:
        raise UnsupportedMethod("discrete-equilibrium central-moment methods are not supported")
    if eq.deviation_only:
        # the truncated delta equilibrium differs from relaxing towards kappa_eq of the full
        # Maxwellian by terms above second order in u; only the default (full) form is built
        raise UnsupportedMethod("central-moment methods with delta_equilibrium=True are not supported")
    polys = list(method.moments)
    monos = set()
    for p in polys:
        for e in p.terms:
            monos.add(tuple(e))
    if len(monos) != Q:
        raise UnsupportedMethod("central-moment polynomials that need de-aliasing are not supported")
    tab = np.zeros(_CENTRAL_SIZE)
    EQ, R0, COEF, PROW, PINV, L0, LU, FMODE = 0, 540, 541, 568, 568 + 729, 568 + 1458, 568 + 1458 + 27, \
        568 + 1458 + 108
    GAUSS, SHEAR_ON, SHEAR = FMODE + 1, FMODE + 2, FMODE + 3
    # kappa_eq - gamma as a polynomial in (drho, u): rho -> 1 + drho
    full = eq
    rho = Poly.const(EQ_VARS, 1) + Poly.var(EQ_VARS, "drho")
    for idx, e in enumerate(_CANON):
        if D == 2 and e[2] != 0:
            continue
        if e not in monos:
            continue
        p = full.monomial_central_moment(e).subs({"rho": rho}) - _gamma(e)
        used = set()
        for t, um in enumerate(_UMONO):
            for k, dr in enumerate((0, 1)):
                ex = (0, dr) + um
                c = p.coeff(ex)
                tab[EQ + 20 * idx + 10 * k + t] = float(c)
                used.add(ex)
        rest = [ex for ex in p.terms if ex not in used]
        if rest:
            raise UnsupportedMethod(f"central-moment equilibrium term outside the kernel's basis: {rest}")
    # polynomial matrix in canonical monomial indexing, rates with a base rate + rank-one corrections
    P = np.zeros((Q, 27))
    for k, p in enumerate(polys):
        for e, c in p.terms.items():
            P[k, _CANON.index(tuple(e))] = float(c)
    cols = [_CANON.index(e) for e in sorted(monos, key=_CANON.index)]
    Psub = P[:, cols]
    Pinv_sub = np.linalg.inv(Psub)
    Pinv = np.zeros((27, Q))
    Pinv[cols, :] = Pinv_sub
    rates = np.array([bind_rate(r, bindings) for r in method.relaxation_rates])
    uniq, counts = np.unique(rates, return_counts=True)
    r0 = float(uniq[np.argmax(counts)])
    tab[R0] = r0
    n = 0
    for k in range(Q):
        if rates[k] != r0:
            tab[COEF + n] = rates[k] - r0
            tab[PROW + 27 * n: PROW + 27 * (n + 1)] = P[k]
            tab[PINV + 27 * n: PINV + 27 * (n + 1)] = Pinv[:, k]
            n += 1
    fm = method.force_model
    if fm is not None:
        t0, tu = fm.population_terms(method)
        tab[L0:L0 + Q] = [float(v) for v in t0]
        for i in range(Q):
            for a in range(D):
                tab[LU + 3 * i + a] = float(tu[i][a])
        tab[FMODE] = 2.0 if type(fm).__name__ == "Guo" else 1.0
    if Q == 19 and any(all(x > 0 for x in e) for e in monos):
        # the kernel's inverse transform sets the raw moments of these monomials to zero (they vanish on D3Q19)
        raise UnsupportedMethod("a D3Q19 central-moment set cannot hold monomials with three non-zero exponents")
    # closed forms the kernel may use instead of the general terms (LBM355_FLAG_GAUSSIAN_EQUILIBRIUM,
    # LBM355_FLAG_SHEAR_PROJECTOR), each checked here against the table it replaces
    eqb = tab[EQ:EQ + 540].reshape(27, 20)
    present = np.array([(e in monos) and not (D == 2 and e[2] != 0) for e in _CANON])
    # 3-D stencils whose monomials form a downward-closed set (every monomial below a used one is used: D3Q27,
    # D3Q19), so that the raw moments the backward transform uses depend on used central moments only
    closed_set = D == 3 and all(tuple(a - (i == k) for i, a in enumerate(e)) in monos
                                for e in monos for k in range(3) if e[k] > 0)
    tab[GAUSS] = _gaussian_mode(eqb, present) if closed_set else 0.0
    s = None if not closed_set else _central_shear_scale(tab[COEF:COEF + 27], tab[PROW:PROW + 729].reshape(27, 27),
                             tab[PINV:PINV + 729].reshape(27, 27), n, r0, D)
    if s is not None:
        tab[SHEAR_ON], tab[SHEAR] = 1.0, s
    return tab.astype(dtype), n


def _gaussian_central_eq(e, at_rest):
    """kappa_eq - gamma of monomial e as the table's 20 coefficients ((1, drho) x _UMONO): for the full continuous
    Maxwellian, drho gamma_e (its central moments are rho gamma_e); for the incompressible one (a rest Gaussian of
    weight rho_0 = 1 centred at u plus one of weight drho centred at 0), drho prod_a g(e_a, u_a) with
    g(0) = 1, g(1) = -u, g(2) = 1/3 + u^2, truncated at second order in u."""
    row = np.zeros(20)
    if not at_rest:
        row[10] = float(_gamma(e))
        return row
    ones = [a for a in range(3) if e[a] == 1]
    twos = [a for a in range(3) if e[a] == 2]
    if len(ones) >= 3:
        return row
    lin = [0, 0, 0]
    for a in ones:
        lin[a] += 1
    sign = -1.0 if len(ones) == 1 else 1.0
    third = Fraction(1, 3)
    terms = {tuple(lin): sign * float(third ** len(twos))}
    if not ones:
        for a in twos:
            q = [0, 0, 0]
            q[a] = 2
            terms[tuple(q)] = float(third ** (len(twos) - 1))
    for um, c in terms.items():
        row[10 + _UMONO.index(um)] += c
    return row


def _gaussian_mode(eqb, present):
    """1 when the central-moment equilibrium table is that of the full continuous Maxwellian, 2 when it is the
    incompressible one (_gaussian_central_eq), 0 otherwise (LBM355_FLAG_GAUSSIAN_EQUILIBRIUM)."""
    for mode, at_rest in ((1.0, False), (2.0, True)):
        want = np.array([_gaussian_central_eq(e, at_rest) if present[i] else np.zeros(20)
                         for i, e in enumerate(_CANON)])
        if np.array_equal(eqb, want):
            return mode
    return 0.0


def central_closed_forms(tab):
    """(Gaussian equilibrium, closed-form rank-one sum) markers of a central-moment table (central_table)."""
    if tab.size != _CENTRAL_SIZE:
        return False, False
    base = _CENTRAL_SIZE - 3
    return bool(tab[base] in (1.0, 2.0)), bool(tab[base + 1] == 1.0)


_CENTRAL_SIZE = 540 + 1 + 27 + 1458 + 27 + 81 + 1 + 3


def _central_shear_scale(coef, prow, pinv, n, r0, D):
    """s when the central-moment table's rank-one sum, sum_k coef_k pinv_k prow_k^T (27 x 27, monomial
    coordinates), equals -r0 on the conserved monomials (000, 100, 010, 001: relaxed with rate 0) plus s times the
    deviatoric projector on the second-order monomials (the shear moments relaxed with r0 + s); None otherwise
    (3-D only; checked to 1e-12)."""
    if D != 3 or n == 0:
        return None
    R = np.zeros((27, 27))
    for k in range(n):
        R += coef[k] * np.outer(pinv[k], prow[k])
    idx = {e: i for i, e in enumerate(_CANON)}
    C = np.zeros((27, 27))
    for e in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)):
        C[idx[e], idx[e]] = -r0
    Dv = np.zeros((27, 27))
    diag = [idx[(2, 0, 0)], idx[(0, 2, 0)], idx[(0, 0, 2)]]
    for a in diag:
        for b in diag:
            Dv[a, b] = (1.0 if a == b else 0.0) - 1.0 / 3.0
    for e in ((1, 1, 0), (1, 0, 1), (0, 1, 1)):
        Dv[idx[e], idx[e]] = 1.0
    rest = R - C
    s = float((rest * Dv).sum() / (Dv * Dv).sum())
    if np.abs(rest - s * Dv).max() > 1e-12 * max(1.0, np.abs(R).max()):
        return None
    return s


def fixed_density_table(method):
    """(Q, 8) FP64 table of the storage-form equilibrium per direction, A, B, Qd[3], Qo[3] (xy, xz, yz), for the
    FixedDensity boundary (its symmetric part: A + B drho + lam (Qd.u^2 + Qo.uu)).  Moment family only."""
    if family_of(method) != FAMILY_MOMENT:
        raise UnsupportedMethod("FixedDensity on the GPU path needs a moment-family method (SRT, TRT, MRT)")
    feq, _, _ = _storage_equilibrium(method)
    tab = np.zeros((method.stencil.Q, 8))
    for i, p in enumerate(feq):
        A, B, _, qd, qo = _split_equilibrium(p, method.compressible)
        tab[i] = [float(A), float(B)] + [float(v) for v in qd] + [float(v) for v in qo]
    return tab


def method_table(method, bindings, dtype=np.float64, corrections=None, inputs=False):
    """``inputs``: the sweep reads density / velocity from input fields (moment family only)."""
    fam = family_of(method)
    if corrections is not None and fam != FAMILY_CUMULANT:
        raise ValueError("Galilean / fourth-order corrections are only defined for D3Q27 cumulant methods.")
    if inputs and fam != FAMILY_MOMENT:
        raise UnsupportedMethod("density / velocity inputs are implemented for the moment family")
    if fam == FAMILY_MOMENT:
        return moment_table(method, bindings, dtype, inputs=inputs)
    if fam == FAMILY_CUMULANT:
        return cumulant_table(method, bindings, dtype, corrections)
    return central_table(method, bindings, dtype)
