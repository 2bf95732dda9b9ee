"""CPU FP64 oracle for the fused stream-collide sweep -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module; the product path (``lbmpy_amd``) never does.

It restates lbmpy's sweep generically, in the matrix / generating-function
form in which lbmpy derives it, vectorised over cells with numpy:

* density/velocity: ``methods/conservedquantitycomputation.py:218-249``;
* moment methods (SRT/TRT in population space, MRT in raw-moment space):
  ``methods/momentbased/momentbasedmethod.py:320-416`` --
  f* = f + M^-1 D (m_eq - M f) + force, or m* = m + D (m_eq - m) + force_m;
* central-moment methods: ``methods/momentbased/centralmomentbasedmethod.py:305-394``
  with the binomial shift of ``moment_transforms/centralmomenttransforms.py:168-483``;
* cumulant methods: ``methods/cumulantbased/cumulantbasedmethod.py:296-402`` with
  the cumulant generating function of ``moment_transforms/cumulanttransforms.py:46-220``
  evaluated by the Leibniz recursion on K' = C' K (no hand-expanded formulas);
* D3Q27 cumulant corrections: Galilean (``methods/cumulantbased/galilean_correction.py:21-91``) and
  fourth-order with limited third-order rates (``methods/cumulantbased/fourth_order_correction.py:21-237``),
  added to the post-collision polynomial cumulants in lbmpy's rho-scaled form;
* force models: ``forcemodels.py:179-316`` (Simple, Luo, Guo, CentralMoment);
* streaming: pull ``fieldaccess.py:66-75`` and AA ``fieldaccess.py:136-157`` on
  arrays with one ghost layer, periodic ghost copies (``lbstep.py:148-151``)
  and the NoSlip index-list boundary kernel
  (``boundaries/boundaryconditions.py:104-139``, ``boundaries/boundaryhandling.py:159-199``)
  executed between steps exactly as ``LatticeBoltzmannStep.time_step`` orders
  them (``lbstep.py:271-282``).

The method description (stencil, moment polynomials, relaxation rates,
equilibrium values, weights) is taken from ``lbmpy_amd.methods`` -- the same
numeric LBMethod the kernels are parameterised from; everything downstream of
it is computed here independently.
"""
from fractions import Fraction
from itertools import product
from math import comb

import numpy as np

from lbmpy_amd.moments import moment_matrix, evaluate_moment, monomial_sort_key
from lbmpy_amd.equilibrium import EQ_VARS


# ------------------------------------------------------------------------ helpers

def _rate(r, bindings):
    if hasattr(r, "free_symbols"):
        return float(r.subs({s: bindings[s.name] for s in r.free_symbols}))
    return float(r)


def _eval_eq_poly(p, rho, drho, u):
    """Evaluate a polynomial in (rho, drho, u0, u1, u2) on cell arrays."""
    vals = {"rho": rho, "drho": drho, "u0": u[0], "u1": u[1] if len(u) > 1 else 0.0,
            "u2": u[2] if len(u) > 2 else 0.0}
    out = np.zeros_like(rho)
    for e, c in p.terms.items():
        t = float(c)
        term = np.full_like(rho, t)
        for v, k in zip(EQ_VARS, e):
            if k:
                term = term * vals[v] ** k
        out = out + term
    return out


def _mono_exponents(polys, dim):
    monos = set()
    for p in polys:
        for e in p.terms:
            monos.add(tuple(e[:dim]) + (0,) * (3 - dim))
    return sorted(monos, key=monomial_sort_key)


class OracleCollision:
    """Collision operator of one method on arrays shaped (Q, ncells)."""

    #: the constants of lbmpy's Smagorinsky rate (turbulence_models.py:62-71): Pi = sqrt(2 Pi_neq:Pi_neq) / rho
    #: and omega = 2 / (tau0 + sqrt(18 Cs^2 Pi + tau0^2)); class attributes so that a test can plant a mistake
    FROBENIUS_FACTOR = 2.0
    SGS_SQRT_FACTOR = 18.0

    def __init__(self, method, bindings=None, smagorinsky=None, closure=None, galilean=False, fourth_order=None,
                 inputs=None):
        """``inputs``: None or (rho_in (n,) or None, u_in (D, n)), the density / velocity input fields over the cells
        of every call (creationfunctions.py:655-680): the equilibrium and force terms take them in place of the
        values computed from f, the macroscopic outputs stay those of f;
        ``smagorinsky``: Smagorinsky constant Cs (turbulence_models.py:25-84) or None;
        ``closure``: callable (pi_neq (n, 6), rho (n,), omega_shear) -> shear rate (n,), or None;
        ``galilean``: the D3Q27 cumulant Galilean correction (galilean_correction.py:21-91);
        ``fourth_order``: None or (limiter, shear rate, bulk rate) for the fourth-order correction
        (fourth_order_correction.py:21-237); the method then carries the rate symbols corr_rr_0..9
        (creationfunctions.py:835-846), bound here as lbmpy's collision rule assigns them."""
        self.m = method
        self.inputs = inputs
        self.smagorinsky = smagorinsky
        self.closure = closure
        self.galilean = bool(galilean)
        self.fourth = fourth_order
        self.bind = dict(bindings or {})
        if fourth_order is not None:
            lim, shear, bulk = fourth_order
            w1, w2 = _rate(shear, self.bind), _rate(bulk, self.bind)
            self.fourth_ab = _fourth_order_ab(w1, w2)
            self.fourth_w = _fourth_order_omegas(w1, w2)
            for i, v in enumerate((w1, w2) + self.fourth_w + (1.0,) * 5):
                self.bind[f"corr_rr_{i}"] = v
        self.stencil = method.stencil
        self.Q = method.stencil.Q
        self.D = method.stencil.D
        self.c = np.array(method.stencil.stencil_entries, dtype=np.float64)          # (Q, D)
        self.w = np.array([float(w) for w in method.weights])
        self.rates = np.array([_rate(r, self.bind) for r in method.relaxation_rates])
        self.polys = list(method.relaxation_rate_dict.keys())
        fm = method.force_model
        self.F = np.array(fm.force if fm is not None else (0.0,) * self.D, dtype=np.float64)
        self.fm = fm
        self.cqc = method.conserved_quantity_computation
        self.eq = method.equilibrium_distribution
        if method.kind == "moment":
            self._setup_moment()
        else:
            self._setup_central()

    # ---- density / velocity (CQC) -----------------------------------------------------
    def macroscopic(self, f):
        s = f.sum(axis=0)
        zc = self.m.zero_centered_pdfs
        rho = 1.0 + s if zc else s
        drho = s if zc else s - 1.0
        j = self.c.T @ f                                           # (D, n)
        ref = rho if self.m.compressible else np.ones_like(rho)
        u = j / ref
        u_eq = u.copy()
        u_out = u.copy()
        if self.fm is not None:
            shift = self.F[:, None] / (2.0 * ref[None, :])
            if self.fm.equilibrium_velocity_shift:
                u_eq = u + shift
            u_out = u + shift
        if self.inputs is not None:
            rho_in, u_in = self.inputs
            u_eq = np.array(u_in, dtype=np.float64).reshape(self.D, -1)
            if rho_in is not None:
                rho = np.array(rho_in, dtype=np.float64).reshape(-1)
                drho = rho - 1.0
        return rho, drho, u_eq, u_out

    def _luo(self, u):
        """Luo term w_i [F.(c-u)/cs2 + (c.F)(c.u)/cs2^2] per direction (forcemodels.py:235-256)."""
        cs2 = 1.0 / 3.0
        cF = self.c @ self.F                                       # (Q,)
        cu = self.c @ u                                            # (Q, n)
        Fu = self.F @ u                                            # (n,)
        return self.w[:, None] * ((cF[:, None] - Fu[None, :]) / cs2 + cF[:, None] * cu / cs2 ** 2)

    def _simple(self, n):
        cs2 = 1.0 / 3.0
        return np.repeat((self.w * (self.c @ self.F) / cs2)[:, None], n, axis=1)

    def _pop_force_vector(self, u):
        name = type(self.fm).__name__
        if name == "Luo" or name == "Guo":
            return self._luo(u)
        if name == "Simple":
            return self._simple(u.shape[1])
        raise NotImplementedError(name)

    # ---- moment methods ---------------------------------------------------------------
    def _setup_moment(self):
        M = moment_matrix(self.polys, self.stencil)
        self.M = np.array([[float(v) for v in r] for r in M])
        self.Minv = np.linalg.inv(self.M)
        self.Dm = np.diag(self.rates)
        self.meq_polys = list(self.m.moment_equilibrium_values)
        self.raw_space = self.m.collision_space.name == "RAW_MOMENTS"
        zc = self.m.zero_centered_pdfs
        self.bg = None
        if zc and not self.eq.deviation_only:
            self.bg = np.array([float(self.eq.background_moment(p)) for p in self.polys])

    def _collide_moment(self, f):
        rho, drho, u, u_out = self.macroscopic(f)
        n = f.shape[1]
        meq = np.stack([_eval_eq_poly(p, rho, drho, list(u)) for p in self.meq_polys])
        if self.raw_space:
            m = self.M @ f
            if self.bg is not None:
                m = m + self.bg[:, None]
            mpost = m + self.Dm @ (meq - m)
            if self.fm is not None:
                name = type(self.fm).__name__
                fvec = self._pop_force_vector(u)
                if name == "Guo":
                    mpost = mpost + (np.eye(self.Q) - 0.5 * self.Dm) @ (self.M @ fvec)
                else:
                    mpost = mpost + self.M @ fvec
            if self.bg is not None:
                mpost = mpost - self.bg[:, None]
            fpost = self.Minv @ mpost
        else:
            fpost = f + self.Minv @ (self.Dm @ (meq - self.M @ f))
            if self.fm is not None:
                name = type(self.fm).__name__
                fvec = self._pop_force_vector(u)
                if name == "Guo":
                    if len(set(self.rates.tolist())) == 1:
                        fvec = (1.0 - self.rates[0] / 2.0) * fvec
                    else:
                        fvec = self.Minv @ ((np.eye(self.Q) - 0.5 * self.Dm) @ (self.M @ fvec))
                fpost = fpost + fvec
        if self.inputs is not None:
            # the density output is the populations' own, as lbmpy's output equations read it
            s = f.sum(axis=0)
            rho = 1.0 + s if self.m.zero_centered_pdfs else s
        return fpost, rho, u_out

    # ---- central-moment and cumulant methods ------------------------------------------
    def _setup_central(self):
        if self.inputs is not None:
            raise NotImplementedError("density / velocity inputs are restated for the moment family")
        D = self.D
        self.monos = _mono_exponents(self.polys, D)
        if len(self.monos) != self.Q:
            raise NotImplementedError("de-aliasing of central-moment polynomials is not restated")
        Mm = moment_matrix(self.monos, self.stencil)
        self.Mmono = np.array([[float(v) for v in r] for r in Mm])
        self.Mmono_inv = np.linalg.inv(self.Mmono)
        idx = {e: i for i, e in enumerate(self.monos)}
        self.midx = idx
        P = np.zeros((self.Q, self.Q))
        for i, p in enumerate(self.polys):
            for e, cf in p.terms.items():
                P[i, idx[tuple(e[:D]) + (0,) * (3 - D)]] = float(cf)
        self.P = P
        self.Pinv = np.linalg.inv(P)
        self.bg = None
        if self.m.zero_centered_pdfs:
            if self.m.kind == "cumulant" or not self.eq.deviation_only:
                self.bg = np.array([_background_monomial(e, D) for e in self.monos])
        # binomial shift pairs: kappa_e = sum_{r<=e} prod C(e_a,r_a) s_a^(e_a-r_a) m_r
        self.shift_pairs = []
        for e in self.monos:
            for r in product(*(range(k + 1) for k in e)):
                self.shift_pairs.append((idx[e], idx[r], tuple(comb(e[a], r[a]) for a in range(3)),
                                         tuple(e[a] - r[a] for a in range(3))))
        if self.m.kind == "central_moment":
            self.keq_polys = list(self.m.moment_equilibrium_values)
        else:
            self.ceq_polys = list(self.m.cumulant_equilibrium_values)

    def _shift(self, m, s):
        """Binomial shift of monomial moments by velocity s (s=-u: raw->central, s=+u: central->raw)."""
        out = np.zeros_like(m)
        for i, r, binom, pw in self.shift_pairs:
            coef = float(binom[0] * binom[1] * binom[2])
            t = coef * m[r]
            for a in range(3):
                if pw[a]:
                    t = t * s[a] ** pw[a]
            out[i] = out[i] + t
        return out

    def _first_order_indices(self):
        out = []
        for a in range(self.D):
            e = [0, 0, 0]
            e[a] = 1
            out.append(self.midx[tuple(e)])
        return out

    def _collide_central(self, f):
        rho, drho, u, u_out = self.macroscopic(f)
        n = f.shape[1]
        u3 = [u[a] if a < self.D else np.zeros(n) for a in range(3)]
        m = self.Mmono @ f
        if self.bg is not None:
            m = m + self.bg[:, None]
        kappa = self._shift(m, [-x for x in u3])
        fo = self._first_order_indices()
        if self.m.kind == "central_moment":
            K = self.P @ kappa
            keq = np.stack([_eval_eq_poly(p, rho, drho, list(u)) for p in self.keq_polys])
            Kp = K + np.diag(self.rates) @ (keq - K)
            if self.fm is not None:
                name = type(self.fm).__name__
                fvec = self._pop_force_vector(u)
                # central moments of the force vector about u: polynomial central moments of fvec
                fk = self.P @ self._shift(self.Mmono @ fvec, [-x for x in u3])
                if name == "Guo":
                    fk = (np.eye(self.Q) - 0.5 * np.diag(self.rates)) @ fk
                Kp = Kp + fk
            kpost = self.Pinv @ Kp
        else:
            if self.fm is not None:
                for a, i in enumerate(fo):
                    kappa[i] = kappa[i] + 0.5 * self.F[a]
            c = self._cumulants_from_central(kappa, rho, u3)
            C = self.P @ c
            ceq = np.zeros_like(C)
            for i, p in enumerate(self.ceq_polys):
                if p is not None:
                    ceq[i] = _eval_eq_poly(p, rho, drho, list(u))
            rates = self.rates
            for i, p in enumerate(self.ceq_polys):
                if p is None:
                    assert rates[i] == 0.0, "the zeroth cumulant must not be relaxed"
            rates = np.repeat(rates[:, None], n, axis=1)
            if self.fourth is not None:
                rates = self._limited_rates(c, rho, rates)
            if self.smagorinsky is not None:
                rates = self._smagorinsky_rates(f, rho, drho, u, rates)
            if self.closure is not None:
                rates = self._closure_rates(f, rho, drho, u, rates)
            Cp = C + rates * (ceq - C)
            if self.galilean:
                self._add_galilean_correction(Cp, c, rho, u3, rates)
            if self.fourth is not None:
                self._add_fourth_order_correction(Cp, c, rho, rates)
            cpost = self.Pinv @ Cp
            kpost = self._central_from_cumulants(cpost, rho, u3, kappa[self.midx[(0, 0, 0)]])
            if self.fm is not None:
                for a, i in enumerate(fo):
                    kpost[i] = kpost[i] + 0.5 * self.F[a]
        mpost = self._shift(kpost, u3)
        if self.bg is not None:
            mpost = mpost - self.bg[:, None]
        return self.Mmono_inv @ mpost, rho, u_out

    def _pi_neq(self, f, rho, drho, u):
        """Pi_neq_ab = sum_i c_ia c_ib (f_i - f_eq_i), f_eq the method's equilibrium populations at (rho, u)
        (utils.py:4-13 as used by turbulence_models.py:59-64); rows xx, yy, zz, xy, xz, yz (3-D)."""
        feq = equilibrium_pdfs_from_eq_velocity(self, rho, u)
        fneq = f - feq
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)] if self.D == 3 else [(0, 0), (1, 1), (0, 1)]
        return np.stack([(self.c[:, a] * self.c[:, b]) @ fneq for a, b in pairs])

    def _smagorinsky_rates(self, f, rho, drho, u, rates):
        """Replace the shear rate by 2 / (tau0 + sqrt(18 Cs^2 Pi + tau0^2)) with
        Pi = sqrt(2 sum_ab Pi_neq_ab^2) / rho_ref, Pi_neq = sum_i c_ia c_ib (f_i - f_eq_i) and
        f_eq the method's equilibrium populations at (rho, u) (turbulence_models.py:59-71,
        utils.py:4-13)."""
        from lbmpy_amd.moments import is_shear_moment
        D = self.D
        p = self._pi_neq(f, rho, drho, u)
        diag = p[:D]
        off = p[D:]
        pi2 = (diag ** 2).sum(0) + 2.0 * (off ** 2).sum(0)
        ref = rho if self.m.compressible else np.ones_like(rho)
        Pi = np.sqrt(self.FROBENIUS_FACTOR * pi2) / ref
        shear = [i for i, p_ in enumerate(self.polys) if is_shear_moment(p_, D)]
        omega_s = self.rates[shear[0]]
        tau0 = 1.0 / omega_s
        w = 2.0 / (tau0 + np.sqrt(self.SGS_SQRT_FACTOR * self.smagorinsky ** 2 * Pi + tau0 ** 2))
        out = rates.copy()
        for i in shear:
            out[i] = w
        return out

    def _closure_rates(self, f, rho, drho, u, rates):
        """Shear rate per cell from ``self.closure(pi_neq (n, 6), rho (n,), omega_shear)`` (a learned closure)."""
        from lbmpy_amd.moments import is_shear_moment
        p = self._pi_neq(f, rho, drho, u)
        shear = [i for i, p_ in enumerate(self.polys) if is_shear_moment(p_, self.D)]
        w = self.closure(p.T, rho, self.rates[shear[0]])
        out = rates.copy()
        for i in shear:
            out[i] = w
        return out

    # ---- cumulant corrections ------------------------------------------------------------
    def _poly_row(self, terms):
        """Index of the polynomial with the given {exponent: coefficient} terms in the method's list."""
        want = {tuple(e): Fraction(v) for e, v in terms.items()}
        for i, p in enumerate(self.polys):
            if {tuple(e[:3]): Fraction(v) for e, v in p.terms.items()} == want:
                return i
        raise ValueError(f"polynomial {want} is not among the method's cumulants")

    def _c(self, c, e):
        return c[self.midx[e]]

    def _velocity_gradients(self, c, rho, w1, w2):
        """D_x, D_y, D_z from the pre-collision second-order cumulants (galilean_correction.py:78-84,
        fourth_order_correction.py:137-146)."""
        cxx, cyy, czz = self._c(c, (2, 0, 0)), self._c(c, (0, 2, 0)), self._c(c, (0, 0, 2))
        dx = -w1 / (2 * rho) * (2 * cxx - cyy - czz) - w2 / (2 * rho) * (cxx + cyy + czz - rho)
        dy = dx + 3 * w1 / (2 * rho) * (cxx - cyy)
        dz = dx + 3 * w1 / (2 * rho) * (cxx - czz)
        return dx, dy, dz

    def _add_galilean_correction(self, Cp, c, rho, u3, rates):
        """Post-collision x2-y2, x2-z2, x2+y2+z2 plus the terms of galilean_correction.py:86-91."""
        i1 = self._poly_row({(2, 0, 0): 1, (0, 2, 0): -1})
        i2 = self._poly_row({(2, 0, 0): 1, (0, 0, 2): -1})
        i3 = self._poly_row({(2, 0, 0): 1, (0, 2, 0): 1, (0, 0, 2): 1})
        w1, w2 = rates[i1], rates[i3]
        if not np.array_equal(w1, rates[i2]):
            raise ValueError("Cumulants (x^2 - y^2) and (x^2 - z^2) must have the same relaxation rate")
        dx, dy, dz = self._velocity_gradients(c, rho, w1, w2)
        ux2, uy2, uz2 = u3[0] ** 2, u3[1] ** 2, u3[2] ** 2
        Cp[i1] = Cp[i1] - 3 * rho * (1 - w1 / 2) * (ux2 * dx - uy2 * dy)
        Cp[i2] = Cp[i2] - 3 * rho * (1 - w1 / 2) * (ux2 * dx - uz2 * dz)
        Cp[i3] = Cp[i3] - 3 * rho * (1 - w2 / 2) * (ux2 * dx + uy2 * dy + uz2 * dz)

    def _limited_rates(self, c, rho, rates):
        """corr_rr_2..4 per cell: w + (1 - w) |C| / (rho lim + |C|) with C = c_xyy + c_xzz, c_xyy - c_xzz,
        c_xyz (fourth_order_correction.py:185-209); every polynomial relaxed with that symbol takes it."""
        lim = float(self.fourth[0])
        sums = (self._c(c, (1, 2, 0)) + self._c(c, (1, 0, 2)), self._c(c, (1, 2, 0)) - self._c(c, (1, 0, 2)),
                self._c(c, (1, 1, 1)))
        out = rates.copy()
        for k, (w, sc) in enumerate(zip(self.fourth_w, sums)):
            lw = w + (1 - w) * np.abs(sc) / (rho * lim + np.abs(sc))
            for i, r in enumerate(self.m.relaxation_rates):
                if getattr(r, "name", None) == f"corr_rr_{2 + k}":
                    out[i] = lw
        return out

    def _add_fourth_order_correction(self, Cp, c, rho, rates):
        """Post-collision fourth-order polynomials plus the terms of fourth_order_correction.py:148-167."""
        x2y2 = (2, 2, 0)
        x2z2 = (2, 0, 2)
        y2z2 = (0, 2, 2)
        rows = [self._poly_row({x2y2: 1, x2z2: -2, y2z2: 1}), self._poly_row({x2y2: 1, x2z2: 1, y2z2: -2}),
                self._poly_row({x2y2: 1, x2z2: 1, y2z2: 1}), self._poly_row({(2, 1, 1): 1}),
                self._poly_row({(1, 2, 1): 1}), self._poly_row({(1, 1, 2): 1})]
        ixy = self._poly_row({(1, 1, 0): 1})
        ibulk = self._poly_row({(2, 0, 0): 1, (0, 2, 0): 1, (0, 0, 2): 1})
        w1, w2 = rates[ixy], rates[ibulk]
        w6, w7, w8 = rates[rows[0]], rates[rows[2]], rates[rows[3]]
        a, b = self.fourth_ab
        dxu, dyv, dzw = self._velocity_gradients(c, rho, w1, w2)
        dxvdyu = -3 * w1 / rho * self._c(c, (1, 1, 0))
        dxwdzu = -3 * w1 / rho * self._c(c, (1, 0, 1))
        dywdzv = -3 * w1 / rho * self._c(c, (0, 1, 1))
        pre = 1 / w1 - 0.5
        Cp[rows[0]] = Cp[rows[0]] + 2 / 3 * pre * w6 * a * rho * (dxu - 2 * dyv + dzw)
        Cp[rows[1]] = Cp[rows[1]] + 2 / 3 * pre * w6 * a * rho * (dxu + dyv - 2 * dzw)
        Cp[rows[2]] = Cp[rows[2]] - 4 / 3 * pre * w7 * a * rho * (dxu + dyv + dzw)
        Cp[rows[3]] = Cp[rows[3]] - 1 / 3 * pre * w8 * b * rho * dywdzv
        Cp[rows[4]] = Cp[rows[4]] - 1 / 3 * pre * w8 * b * rho * dxwdzu
        Cp[rows[5]] = Cp[rows[5]] - 1 / 3 * pre * w8 * b * rho * dxvdyu

    def _leibniz_terms(self, e):
        """Pick j with e_j > 0; return (j, list of (binomial, r)) for r <= e - e_j."""
        j = next(a for a in range(3) if e[a] > 0)
        em = list(e)
        em[j] -= 1
        em = tuple(em)
        terms = []
        for r in product(*(range(k + 1) for k in em)):
            b = comb(em[0], r[0]) * comb(em[1], r[1]) * comb(em[2], r[2])
            rj = list(r)
            rj[j] += 1
            rest = tuple(em[a] - r[a] for a in range(3))
            terms.append((b, tuple(rj), rest))
        return terms

    def _cumulants_from_central(self, kappa, rho, u3):
        """c_e = rho * d^e [Xi.u + log K(Xi)] at 0 (cumulanttransforms.py:157-181)."""
        order = sorted(self.monos, key=lambda e: sum(e))
        L = {}
        k0 = kappa[self.midx[(0, 0, 0)]]
        for e in order:
            if sum(e) == 0:
                continue
            acc = kappa[self.midx[e]].copy()
            for b, rj, rest in self._leibniz_terms(e):
                if rj == e:
                    continue
                acc = acc - b * L[rj] * kappa[self.midx[rest]]
            L[e] = acc / k0
        c = np.zeros_like(kappa)
        for e in self.monos:
            i = self.midx[e]
            if sum(e) == 0:
                c[i] = rho * np.log(k0)
            elif sum(e) == 1:
                a = e.index(1)
                c[i] = rho * (u3[a] + L[e])
            else:
                c[i] = rho * L[e]
        return c

    def _central_from_cumulants(self, cpost, rho, u3, k0):
        """kappa_e = d^e exp(C(Xi) - Xi.u) at 0 with c -> c/rho (cumulanttransforms.py:183-214);
        exp(c_000/rho) = exp(log kappa_000) = kappa_000."""
        order = sorted(self.monos, key=lambda e: sum(e))
        L = {}
        for e in self.monos:
            i = self.midx[e]
            if sum(e) == 1:
                a = e.index(1)
                L[e] = cpost[i] / rho - u3[a]
            elif sum(e) > 1:
                L[e] = cpost[i] / rho
        K = {(0, 0, 0): k0}
        for e in order:
            if sum(e) == 0:
                continue
            acc = np.zeros_like(k0)
            for b, rj, rest in self._leibniz_terms(e):
                acc = acc + b * L[rj] * K[rest]
            K[e] = acc
        out = np.zeros_like(cpost)
        for e in self.monos:
            out[self.midx[e]] = K[e]
        return out

    def __call__(self, f):
        if self.m.kind == "moment":
            return self._collide_moment(f)
        return self._collide_central(f)


def _fourth_order_ab(w1, w2):
    """a, b of Geier et al. 2017 eq. 115-116 (fourth_order_correction.py:70-94)."""
    den = (w1 - w2) * (w2 * (2 + 3 * w1) - 8 * w1)
    a = (4 * w1 ** 2 + 2 * w1 * w2 * (w1 - 6) + w2 ** 2 * (w1 * (10 - 3 * w1) - 4)) / den
    b = (4 * w1 * w2 * (9 * w1 - 16) - 4 * w1 ** 2 - 2 * w2 ** 2 * (2 + 9 * w1 * (w1 - 2))) / (3 * den)
    return a, b


def _fourth_order_omegas(w1, w2):
    """Unlimited omega_3, omega_4, omega_5 of Geier et al. 2017 eq. 112-114
    (fourth_order_correction.py:127-139)."""
    o3 = 8 * (w1 - 2) * (w2 * (3 * w1 - 1) - 5 * w1) / (8 * (5 - 2 * w1) * w1 + w2 * (8 + w1 * (9 * w1 - 26)))
    o4 = 8 * (w1 - 2) * (w1 + w2 * (3 * w1 - 7)) / (w2 * (56 - 42 * w1 + 9 * w1 ** 2) - 8 * w1)
    o5 = 24 * (w1 - 2) * (4 * w1 ** 2 + w1 * w2 * (18 - 13 * w1) + w2 ** 2 * (2 + w1 * (6 * w1 - 11))) / \
        (16 * w1 ** 2 * (w1 - 6) - 2 * w1 * w2 * (216 + 5 * w1 * (9 * w1 - 46)) +
         w2 ** 2 * (w1 * (3 * w1 - 10) * (15 * w1 - 28) - 48))
    return o3, o4, o5


def corrections_from_config(lbm_config):
    """OracleCollision keyword arguments for an LBMConfig's cumulant corrections
    (creationfunctions.py:692-707: limiter 1e6 for True, bulk rate 1 when only the shear rate is given)."""
    foc = lbm_config.fourth_order_correction
    rr = list(lbm_config.relaxation_rates)
    fourth = None
    if foc:
        fourth = (1e6 if foc is True else float(foc), rr[0], rr[1] if len(rr) > 1 else 1)
    return {"galilean": bool(lbm_config.galilean_correction), "fourth_order": fourth}


def _background_monomial(e, dim):
    """Raw monomial moment of the rest state rho=1, u=0 (Gaussian with cs2=1/3)."""
    v = 1.0
    for a in range(dim):
        k = e[a]
        if k % 2:
            return 0.0
        v *= (1.0 / 3.0) ** (k // 2) * (1 if k < 4 else 3)
    return v


# ------------------------------------------------------------------------ streaming driver

# ------------------------------------------------------------------------ streaming accessors

# lbmpy's 'lehmann' orderings (stencils.py:146-149, 214-218, 242-248); D3Q19 is the first 19 of D3Q27's
_LEHMANN_2D = ((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1))
_LEHMANN_3D = ((0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
               (1, 1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, -1, -1), (1, -1, 0), (-1, 1, 0),
               (1, 0, -1), (-1, 0, 1), (0, 1, -1), (0, -1, 1), (1, 1, 1), (-1, -1, -1), (1, 1, -1), (-1, -1, 1),
               (1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1))

INPLACE_PATTERNS = ("aa", "esotwist", "esopull", "esopush")


def streaming_accessor(c, inv, pattern, timestep):
    """Read and write accesses of one kernel (fieldaccess.py:66-87 pull / push, :136-157 AA,
    :160-316 EsoTwist / EsoPull / EsoPush) for directions c (Q, D) with inverse indices inv:
    (read offsets (Q, D), read slots (Q,), write offsets (Q, D), write slots (Q,)).  timestep: 0 EVEN,
    1 ODD (ignored by pull / push)."""
    c = np.asarray(c, dtype=int)
    Q, D = c.shape
    inv = np.asarray(inv)
    zero = np.zeros_like(c)
    idx = np.arange(Q)
    lehmann = _LEHMANN_2D if D == 2 else _LEHMANN_3D[:Q]
    # EsoPull / EsoPush: "even" directions sit at even positions of the lehmann ordering
    even = np.array([lehmann.index(tuple(v)) % 2 == 0 for v in c])[:, None]
    if pattern == "pull":
        return -c, idx, zero, idx
    if pattern == "push":
        return zero, idx, c, idx
    if pattern == "aa":
        return (zero, idx, zero, inv) if timestep == 0 else (-c, inv, c, idx)
    if pattern == "esotwist":
        up, down = np.maximum(c, 0), np.maximum(-c, 0)
        return (down, idx, up, inv) if timestep == 0 else (down, inv, up, idx)
    if pattern == "esopull":
        r = np.where(even, -c, zero)
        w = np.where(even, zero, c)
        return (r, idx, w, inv) if timestep == 0 else (r, inv, w, idx)
    if pattern == "esopush":
        r = np.where(even, zero, -c)
        w = np.where(even, c, zero)
        return (r, inv, w, idx) if timestep == 0 else (r, idx, w, inv)
    raise ValueError(f"unknown streaming pattern {pattern!r}")


def freeslip_normals(c, cells, dirs):
    """FreeSlip without a given normal (boundaryconditions.py:402-440): for the links (cells[k], dirs[k]) of one
    boundary object, the normal from the wall into the fluid and the reference direction index.  An axis a is a
    wall axis of a link when the fluid cell shifted by c_d's a-component alone is one of the object's boundary
    cells (cells + c_d); the normal is -c_d on wall axes and 0 elsewhere, the reference direction keeps c_d on
    wall axes and reverses it elsewhere.  With no wall axis (a convex corner) the normal is -c_d and the reference
    direction d itself."""
    c = np.asarray(c, dtype=int)
    D = c.shape[1]
    lookup = {tuple(v): i for i, v in enumerate(c.tolist())}
    boundary = {tuple(v) for v in (cells + c[dirs]).tolist()}
    normals = np.zeros((len(dirs), D), dtype=int)
    refs = np.zeros(len(dirs), dtype=int)
    for k in range(len(dirs)):
        cd = c[dirs[k]]
        wall = np.array([tuple((cells[k] + np.where(np.arange(D) == a, cd, 0)).tolist()) in boundary
                         for a in range(D)])
        if wall.any():
            normals[k] = np.where(wall, -cd, 0)
            refs[k] = lookup[tuple(np.where(wall, cd, -cd).tolist())]
        else:
            normals[k] = -cd
            refs[k] = dirs[k]
    return normals, refs


class OracleDomain:
    """A domain with one ghost layer, stepping the way ``LatticeBoltzmannStep`` does.

    ``faces`` maps axis -> 'periodic' | 'noslip' | 'ghost' (both sides of the axis).
    Arrays are numpy float64 with shape (Q, nz+2, ny+2, nx+2) for 3-D and
    (Q, ny+2, nx+2) for 2-D (the ``fzyx`` layout)."""

    def __init__(self, method, shape, faces=None, pattern="pull", bindings=None, obstacle=None, boundaries=None,
                 collision=None):
        """``obstacle``: optional bool array [x, y, (z)] over the interior marking NoSlip cells
        (what ``set_boundary(NoSlip(), mask_callback=...)`` flags).  ``boundaries``: optional list of
        (flags, spec) with flags a bool array [x, y, (z)] over the ghost-inclusive domain and spec
        {"kind": "noslip"}, {"kind": "ubb", "velocity": u_w, "density": rho_w or None[, "adapt": True]} or
        {"kind": "freeslip", "normal": n} (n points from the wall into the fluid),
        {"kind": "fixeddensity", "density": rho_w} or {"kind": "outflow", "normal": n} (n points out of the
        fluid).  ``collision``: extra OracleCollision keyword arguments (subgrid model, corrections)."""
        self.boundaries = list(boundaries or [])
        self._outflow_state = {}
        self.coll = OracleCollision(method, bindings, **(collision or {}))
        self.obstacle = None if obstacle is None else np.asarray(obstacle, dtype=bool)
        self.method = method
        self.D = method.stencil.D
        self.Q = method.stencil.Q
        self.shape = tuple(shape)                       # (nx, ny[, nz])
        self.faces = faces or {a: "periodic" for a in range(self.D)}
        self.pattern = pattern
        self.c = np.array(method.stencil.stencil_entries, dtype=int)
        self.inv = np.array(method.stencil.inverse_indices())
        gshape = tuple(n + 2 for n in reversed(self.shape))
        self.src = np.zeros((self.Q,) + gshape)
        self.dst = np.zeros_like(self.src) if pattern in ("pull", "push") else None
        self.aa_parity = 0            # number of executed in-place steps mod 2 (0: next kernel is ODD)
        if pattern not in ("pull", "push") + INPLACE_PATTERNS:
            raise ValueError(f"unknown streaming pattern {pattern!r}")
        self.rho_out = None
        self.u_out = None

    # numpy axis for lattice axis a (arrays are [q, z, y, x])
    def _ax(self, a):
        return self.D - a

    def interior(self, arr=None):
        arr = self.src if arr is None else arr
        sl = (slice(None),) + (slice(1, -1),) * self.D
        return arr[sl]

    def _sl(self, a, idx):
        s = [slice(None)] * (self.D + 1)
        s[self._ax(a)] = idx
        return tuple(s)

    def sync_periodic(self, arr):
        for a in range(self.D):
            if self.faces.get(a) == "periodic":
                arr[self._sl(a, 0)] = arr[self._sl(a, -2)]
                arr[self._sl(a, -1)] = arr[self._sl(a, 1)]

    def wall_links(self):
        """Index list (x, y[, z], dir, kind) in ghost coordinates of fluid cells whose neighbour in
        direction dir is a boundary cell: a NoSlip ghost plane or obstacle cell (kind -1) or a cell of
        ``boundaries[kind]`` (what pystencils' index-list creation yields, boundaries/boundaryhandling.py).
        Neighbours are addressed unwrapped: a boundary plane flags the whole ghost row, including the
        ghost columns of periodic axes, as pystencils' flag field does."""
        if hasattr(self, "_links"):
            return self._links
        n = self.shape
        D = self.D
        gshape = tuple(k + 2 for k in n)
        flags = np.zeros(gshape, dtype=bool)          # [x, y, (z)] ghost-inclusive, NoSlip
        for a in range(D):
            if self.faces.get(a) == "noslip":
                lo = [slice(None)] * D
                hi = [slice(None)] * D
                lo[a], hi[a] = 0, -1
                flags[tuple(lo)] = True
                flags[tuple(hi)] = True
        if self.obstacle is not None:
            flags[(slice(1, -1),) * D] |= self.obstacle
        kind = np.full(gshape, -2, dtype=int)         # -2 fluid, -1 NoSlip, k >= 0 boundaries[k]
        for k in range(len(self.boundaries) - 1, -1, -1):
            kind[self.boundaries[k][0]] = k
        kind[flags] = -1
        # the ghost layers of periodic axes hold the periodic images of the interior cells' flags
        for a in range(D):
            if self.faces.get(a) == "periodic":
                lo = [slice(None)] * D
                hi = [slice(None)] * D
                lo[a], hi[a] = 0, -1
                src_lo = [slice(None)] * D
                src_hi = [slice(None)] * D
                src_lo[a], src_hi[a] = -2, 1
                kind[tuple(lo)] = kind[tuple(src_lo)]
                kind[tuple(hi)] = kind[tuple(src_hi)]
        self.flags = kind > -2
        grids = np.meshgrid(*[np.arange(1, n[a] + 1) for a in range(D)], indexing="ij")
        cells = np.stack([g.reshape(-1) for g in grids], axis=1)            # (ncells, D)
        cells = cells[~self.flags[tuple(cells[:, a] for a in range(D))]]    # fluid cells only
        rows = []
        for d in range(self.Q):
            nb = cells + self.c[d][None, :]
            kk = kind[tuple(nb[:, a] for a in range(D))]
            hit = kk > -2
            sel = cells[hit]
            rows.append(np.concatenate([sel, np.full((len(sel), 1), d), kk[hit][:, None]], axis=1))
        self._links = np.concatenate(rows) if rows else np.zeros((0, D + 2), dtype=int)
        return self._links

    def _link_neighbours(self, links):
        d = links[:, self.D]
        return [links[:, a] + self.c[d, a] for a in range(self.D)]

    def _accessor(self, pattern, timestep):
        return streaming_accessor(self.c, self.inv, pattern, timestep)

    def _kernel_key(self, prev):
        """(pattern, timestep) of a kernel; "pull", "even", "odd" are the pull and AA kernels."""
        if prev == "pull":
            return ("pull", 2)
        if prev in ("even", "odd"):
            return (self.pattern if self.pattern in INPLACE_PATTERNS else "aa", 0 if prev == "even" else 1)
        return prev

    def _out_addr(self, prev, i, cell):
        """f_out(i) at `cell` after the kernel `prev`: its write accessor (indexing.py:41-46)."""
        pat, t = self._kernel_key(prev)
        _, _, woff, wslot = self._accessor(pat, t)
        coords = [cell[a] + woff[i, a] for a in range(self.D)]
        return (wslot[i],) + tuple(reversed(coords))

    def _in_addr(self, prev, i, cell):
        """f_in(i) at `cell`: the read accessor of the kernel that follows `prev`."""
        pat, t = self._kernel_key(prev)
        roff, rslot, _, _ = self._accessor(pat, t if t == 2 else 1 - t)
        coords = [cell[a] + roff[i, a] for a in range(self.D)]
        return (rslot[i],) + tuple(reversed(coords))

    def apply_links(self, arr, prev):
        """Boundary links between time steps; `prev` is the kernel that ran last ("pull", "even", "odd"):
        NoSlip  f_in(inv d) = f_out(d)                                  (boundaryconditions.py:104-139)
        UBB     f_in(inv d) = f_out(d) - 6 w_d rho (c_d . u_w)          (boundaryconditions.py:724-842)
                rho = sum_i f_out(i) (compressible; + 1 zero-centred), the given wall density, or 1
        FreeSlip f_in(inv d) = f_out(inv(mirror(d))) at x + c_d + n   (boundaryconditions.py:360-487)"""
        links = self.wall_links()
        if len(links) == 0:
            return
        D = self.D
        for k in sorted(set(links[:, D + 1].tolist())):
            sel = links[links[:, D + 1] == k]
            d = sel[:, D]
            cell = [sel[:, a] for a in range(D)]
            spec = {"kind": "noslip"} if k < 0 else self.boundaries[k][1]
            target = self._in_addr(prev, self.inv[d], cell)
            if spec["kind"] == "noslip":
                arr[target] = arr[self._out_addr(prev, d, cell)]
            elif spec["kind"] == "ubb":
                uw = np.asarray(spec["velocity"], dtype=np.float64)
                w = self.coll.w[d]
                if spec.get("density") is not None:
                    rho = np.full(len(d), float(spec["density"]))
                elif self.method.compressible:
                    rho = sum(arr[self._out_addr(prev, np.full_like(d, i), cell)] for i in range(self.Q))
                    rho = rho + (1.0 if self.method.zero_centered_pdfs else 0.0)
                else:
                    rho = np.ones(len(d))
                cu = self.c[d, :D] @ uw[:D]
                if spec.get("adapt") and self.method.force_model is not None:
                    # adapt_velocity_to_force: u_w - F / (2 rho_ref) (boundaryconditions.py:808-812 with
                    # conservedquantitycomputation.py:251-266; rho_ref = rho for compressible methods, else 1)
                    ref = rho if self.method.compressible else np.ones(len(d))
                    cu = cu - (self.c[d, :D] @ self.coll.F[:D]) / (2.0 * ref)
                arr[target] = arr[self._out_addr(prev, d, cell)] - 6.0 * w * cu * rho
            elif spec["kind"] == "freeslip" and spec.get("normal") is None:
                # a normal and a reference direction per link (boundaryconditions.py:402-440)
                cells = np.stack(cell, axis=1)
                nrm, ref = freeslip_normals(self.c[:, :D], cells, d)
                src_cell = [cell[a] + self.c[d, a] + nrm[:, a] for a in range(D)]
                arr[target] = arr[self._out_addr(prev, ref, src_cell)]
            elif spec["kind"] == "freeslip":
                nrm = np.asarray(spec["normal"], dtype=int)
                ax = int(np.nonzero(nrm)[0][0])
                mirrored = self.c[d].copy()
                mirrored[:, ax] *= -1
                mi = np.array([self._dir_index(tuple(v)) for v in mirrored])
                src_cell = [cell[a] + self.c[d, a] + nrm[a] for a in range(D)]
                arr[target] = arr[self._out_addr(prev, self.inv[mi], src_cell)]
            elif spec["kind"] == "simple_outflow":
                # copy along the normal (boundaryconditions.py:844-891): f_in(inv d) = f_out(inv d) at x + c_d - n
                nrm = np.asarray(spec["normal"], dtype=int)
                src = [cell[a] + self.c[d, a] - nrm[a] for a in range(D)]
                arr[target] = arr[self._out_addr(prev, self.inv[d], src)]
            elif spec["kind"] == "outflow":
                # extrapolation outflow (boundaryconditions.py:894-1042) with per-link state (s0, s1)
                nrm = np.asarray(spec["normal"], dtype=int)
                src = [cell[a] + self.c[d, a] - nrm[a] for a in range(D)]
                v = arr[self._out_addr(prev, self.inv[d], src)]
                st = self._outflow_state.get(k)
                if st is None:
                    st = np.stack([v, v])
                c = np.sqrt(1.0 / 3.0)
                inter = st[0] * c + (1.0 - c) * st[1]
                arr[target] = inter
                self._outflow_state[k] = np.stack([v, inter])
            elif spec["kind"] == "fixeddensity":
                arr[target] = self._fixed_density_value(arr, prev, d, cell, float(spec["density"]))
            else:
                raise NotImplementedError(spec["kind"])

    def _fixed_density_value(self, arr, prev, d, cell, rho_w):
        """2 f_eq,sym(d; rho_w, u) - f_out(d) (boundaryconditions.py:1044-1102): u from the fluid cell's
        populations with the density replaced by rho_w; the symmetric part (f_eq(u) + f_eq(-u)) / 2 of the
        method's equilibrium (moment family)."""
        co = self.coll
        if self.method.kind != "moment":
            raise NotImplementedError("the oracle's FixedDensity covers the moment family")
        fo = np.stack([arr[self._out_addr(prev, np.full_like(d, i), cell)] for i in range(self.Q)])
        j = self.c[:, :self.D].T @ fo                                       # (D, n)
        n = len(d)
        rho = np.full(n, rho_w)
        ref = rho if self.method.compressible else np.ones(n)
        u = j / ref
        if co.fm is not None and co.fm.equilibrium_velocity_shift:
            u = u + co.F[:self.D, None] / (2.0 * ref)

        def feq(uu):
            meq = np.stack([_eval_eq_poly(p, rho, rho - 1.0, list(uu)) for p in co.meq_polys])
            if co.bg is not None:
                meq = meq - co.bg[:, None]
            return co.Minv @ meq
        sym = 0.5 * (feq(u) + feq(-u))
        return 2.0 * sym[d, np.arange(n)] - fo[d, np.arange(n)]

    def _dir_index(self, v):
        for i in range(self.Q):
            if tuple(self.c[i, :self.D]) == tuple(v[:self.D]):
                return i
        raise KeyError(v)

    def pull_pre_collision(self):
        """The populations the next pull step collides, after periodicity and the boundary links, as (Q, [z,] y, x),
        without advancing the domain (not for outflow boundaries, whose links keep state)."""
        if self.pattern != "pull":
            raise ValueError("pull_pre_collision is for the pull pattern")
        arr = self.src.copy()
        self.sync_periodic(arr)
        self.noslip_pull(arr)
        return np.stack([self._neighbor_view(arr, i, -self.c[i]) for i in range(self.Q)])

    def noslip_pull(self, arr):
        """Boundary links before a pull step (the previous kernel was a pull sweep)."""
        self.apply_links(arr, "pull")

    def _neighbor_view(self, arr, q, off):
        """arr[q] at (x + off) over the interior."""
        s = [q]
        for a in reversed(range(self.D)):
            lo = 1 + off[a]
            hi = lo + self.shape[a]
            s.append(slice(lo, hi))
        return arr[tuple(s)]

    def _collide_cells(self, fpre):
        shape = fpre.shape[1:]
        fpost, rho, u_out = self.coll(fpre.reshape(self.Q, -1))
        self.rho_out = rho.reshape(shape)
        self.u_out = u_out.reshape((self.D,) + shape)
        return fpost.reshape((self.Q,) + shape)

    def _generic_step(self):
        """push, EsoTwist, EsoPull, EsoPush: periodicity, boundary links of the previous kernel, then the kernel
        through its accessor tables, periodic images of ghost writes folded back (lbstep / communication order)."""
        if self.pattern == "push":
            prev = next_k = ("push", 2)
            arr_in, arr_out = self.src, self.dst
        else:
            t = 1 if self.aa_parity == 0 else 0
            prev, next_k = (self.pattern, 1 - t), (self.pattern, t)
            arr_in = arr_out = self.src
            if any(self.faces.get(a) == "ghost" for a in range(self.D)) and not self.boundaries:
                raise NotImplementedError("the in-place oracle needs periodic or boundary-covered faces")
        self.sync_periodic(arr_in)
        self.apply_links(arr_in, prev)
        roff, rslot, woff, wslot = self._accessor(*next_k)
        fpre = np.stack([self._neighbor_view(arr_in, rslot[i], roff[i]) for i in range(self.Q)])
        fpost = self._collide_cells(fpre)
        for i in range(self.Q):
            self._neighbor_view(arr_out, wslot[i], woff[i])[...] = fpost[i]
        self._fold_writes(arr_out, woff, wslot)
        if self.pattern == "push":
            self.src, self.dst = self.dst, self.src
        else:
            self.aa_parity ^= 1

    def _fold_writes(self, arr, woff, wslot):
        """Move values a kernel wrote into ghost cells of periodic axes to their periodic images."""
        for i in range(self.Q):
            for a in range(self.D):
                if self.faces.get(a) != "periodic" or woff[i, a] == 0:
                    continue
                ax = self._ax(a) - 1
                src_sl = [slice(None)] * self.D
                dst_sl = [slice(None)] * self.D
                src_sl[ax], dst_sl[ax] = (-1, 1) if woff[i, a] == 1 else (0, -2)
                arr[wslot[i]][tuple(dst_sl)] = arr[wslot[i]][tuple(src_sl)]

    def step(self):
        if self.pattern in ("push", "esotwist", "esopull", "esopush"):
            self._generic_step()
            return
        if self.pattern == "pull":
            self.sync_periodic(self.src)
            self.noslip_pull(self.src)
            fpre = np.stack([self._neighbor_view(self.src, i, -self.c[i]) for i in range(self.Q)])
            fpost = self._collide_cells(fpre)
            self.interior(self.dst)[...] = fpost
            self.src, self.dst = self.dst, self.src
        else:
            if any(self.faces.get(a) == "ghost" for a in range(self.D)) and not self.boundaries:
                # populations the ODD step writes into a ghost plane come back only through boundary links
                raise NotImplementedError("the AA oracle needs periodic, NoSlip or boundary-covered faces")
            if self.aa_parity == 0:
                # periodicity(EVEN), boundary(prev=EVEN), then the ODD kernel:
                # read f[x - c_i](inv i), write f[x + c_i](i)
                self.sync_periodic(self.src)
                self.noslip_aa(self.src, prev_even=True)
                fpre = np.stack([self._neighbor_view(self.src, self.inv[i], -self.c[i]) for i in range(self.Q)])
                fpost = self._collide_cells(fpre)
                for i in range(self.Q):
                    self._neighbor_view(self.src, i, self.c[i])[...] = fpost[i]
                self._fold_ghost_writes()
            else:
                # boundary(prev=ODD), then the EVEN kernel: read f[x](i), write f[x](inv i)
                self.noslip_aa(self.src, prev_even=False)
                fpre = self.interior().copy()
                fpost = self._collide_cells(fpre)
                out = self.interior()
                for i in range(self.Q):
                    out[self.inv[i]] = fpost[i]
            self.aa_parity ^= 1

    def noslip_aa(self, arr, prev_even):
        """Boundary links between AA time steps (``advanced_streaming/indexing.py:12-181``): after an
        EVEN step the population leaving x in direction d sits in slot inv(d) of x and is placed where
        the ODD step reads it (slot d of the wall cell); after an ODD step the value written into the
        wall cell (slot d) is returned to slot inv(d) of x, where the EVEN step reads it."""
        self.apply_links(arr, "even" if prev_even else "odd")

    def _fold_ghost_writes(self):
        """Values the ODD kernel wrote into ghost cells belong to the periodic images
        (the in-place periodicity handling of ``advanced_streaming/communication.py``)."""
        for i in range(self.Q):
            ci = self.c[i]
            arr = self.src[i]
            # wrap each periodic axis in turn; edges and corners are handled by the repeated wrapping
            for a in range(self.D):
                if self.faces.get(a) != "periodic":
                    continue
                ax = self._ax(a) - 1
                if ci[a] == 1:
                    src_sl = [slice(None)] * self.D
                    dst_sl = [slice(None)] * self.D
                    src_sl[ax] = -1
                    dst_sl[ax] = 1
                    arr[tuple(dst_sl)] = arr[tuple(src_sl)]
                elif ci[a] == -1:
                    src_sl = [slice(None)] * self.D
                    dst_sl = [slice(None)] * self.D
                    src_sl[ax] = 0
                    dst_sl[ax] = -2
                    arr[tuple(dst_sl)] = arr[tuple(src_sl)]

    # ---- setup / readout ----------------------------------------------------------------
    def set_pdfs_interior(self, f):
        self.interior()[...] = f

    def _last_kernel(self):
        if self.pattern in ("pull", "push"):
            return (self.pattern, 2)
        return (self.pattern, 0 if self.aa_parity == 0 else 1)

    def set_pdfs_natural(self, f):
        """Place post-collision populations f (Q, [nz,] ny, nx) where the kernel that ran last wrote them (pull /
        push, or the EVEN kernel of an in-place pattern before the first step), as lbmpy's setter does with
        ``previous_timestep`` (macroscopic_value_kernels.py:29-45)."""
        _, _, woff, wslot = self._accessor(*self._last_kernel())
        for i in range(self.Q):
            self._neighbor_view(self.src, wslot[i], woff[i])[...] = f[i]
        self._fold_writes(self.src, woff, wslot)

    def pdfs_interior(self):
        return self.interior().copy()

    def pdfs_in_natural_layout(self):
        """Post-collision populations of the last step at each cell (slot i holds f*_i(x)),
        read through the write accessor of the last kernel (macroscopic_value_kernels.py:14-26):
        pull: dst[i][x]; AA after EVEN (or initialisation): f[inv i][x]; AA after ODD: f[i][x + c_i];
        push and the Eso patterns through their accessor tables."""
        if self.pattern in ("push", "esotwist", "esopull", "esopush"):
            _, _, woff, wslot = self._accessor(*self._last_kernel())
            return np.stack([self._neighbor_view(self.src, wslot[i], woff[i]) for i in range(self.Q)])
        if self.pattern != "aa":
            return self.pdfs_interior()
        if self.aa_parity == 0:
            return self.pdfs_interior()[self.inv]
        return np.stack([self._neighbor_view(self.src, i, self.c[i]) for i in range(self.Q)])


def equilibrium_pdfs(method, rho, u, bindings=None):
    """Equilibrium populations for given (rho, u) fields, in the stored format
    (``macroscopic_value_kernels.py:29-45`` via ``conservedquantitycomputation.py:251-272``):
    the equilibrium velocity is u - F/(2 rho_ref) and collision with all rates 1
    maps any state with these moments to the equilibrium (so unbound rate symbols are set to 1)."""
    bind = {name: 1.0 for name in method.relaxation_rate_symbols()}
    bind.update(bindings or {})
    coll = OracleCollision(method, bind)
    shape = rho.shape
    rho_f = rho.reshape(-1).astype(np.float64)
    u_f = u.reshape(u.shape[0], -1).astype(np.float64)
    ref = rho_f if method.compressible else np.ones_like(rho_f)
    if method.force_model is not None:
        u_f = u_f - coll.F[:, None] / (2.0 * ref[None, :])
    drho = rho_f - 1.0
    if method.kind == "moment":
        meq = np.stack([_eval_eq_poly(p, rho_f, drho, list(u_f)) for p in coll.meq_polys])
        if coll.bg is not None:
            meq = meq - coll.bg[:, None]
        feq = coll.Minv @ meq
    else:
        # continuous Maxwellian central moments (rho * Gaussian, truncated as in the method),
        # shifted to raw moments and inverted: the equilibrium of the central/cumulant method
        n = rho_f.shape[0]
        u3 = [u_f[a] if a < coll.D else np.zeros(n) for a in range(3)]
        eq = method.equilibrium_distribution
        if method.kind == "cumulant":
            keq = np.stack([_gauss_central(e, rho_f, coll.D) for e in coll.monos])
        else:
            from lbmpy_amd._poly import moment_poly
            keq = np.stack([_eval_eq_poly(eq.monomial_central_moment(e), rho_f, drho, list(u_f))
                            for e in coll.monos])
        m = coll._shift(keq, u3)
        if coll.bg is not None:
            m = m - coll.bg[:, None]
        feq = coll.Mmono_inv @ m
    return feq.reshape((coll.Q,) + shape)


def strain_rate_tensor(method, f, bindings=None):
    """S = -1.5 (omega_s / rho_ref) Pi(f - f_eq) with f_eq the method's equilibrium populations at the density and
    equilibrium velocity of f (macroscopic_value_kernels.py:67-85); f: natural populations (Q, cells); returns
    (D*D, cells), row-major in (a, b).  ``bindings``: values of symbolic relaxation rates."""
    coll = OracleCollision(method, {s: 1.0 for s in method.relaxation_rate_symbols()})
    rho, _, u_eq, _ = coll.macroscopic(f)
    ref = rho if method.compressible else np.ones_like(rho)
    # equilibrium_pdfs removes F / (2 rho_ref) from the velocity it is given: hand it the equilibrium velocity
    # plus that shift
    u_in = u_eq + (coll.F[:, None] / (2.0 * ref[None, :]) if method.force_model is not None else 0.0)
    feq = equilibrium_pdfs(method, rho, u_in)
    D = method.dim
    fneq = f - feq
    pi = np.stack([(coll.c[:, a] * coll.c[:, b]) @ fneq for a in range(D) for b in range(D)])
    return -1.5 * method_shear_rate(method, bindings) / ref * pi


def method_shear_rate(method, bindings=None):
    """The numeric rate of the method's shear moments (relaxationrates.py:23-46 for a method whose shear
    moments share one rate)."""
    from lbmpy_amd.moments import is_shear_moment
    bind = {s: 1.0 for s in method.relaxation_rate_symbols()}
    bind.update(bindings or {})
    coll = OracleCollision(method, bind)
    rates = {float(coll.rates[i]) for i, p in enumerate(coll.polys) if is_shear_moment(p, method.dim)}
    if len(rates) != 1:
        raise ValueError("the shear moments do not share one rate")
    return rates.pop()


def equilibrium_pdfs_from_eq_velocity(coll, rho, u):
    """Stored-format equilibrium populations of a central-moment/cumulant method at the
    equilibrium velocity u (no force shift applied): Gaussian central moments times rho,
    shifted to raw moments, minus the background, M^-1."""
    n = rho.shape[0]
    u3 = [u[a] if a < coll.D else np.zeros(n) for a in range(3)]
    keq = np.stack([_gauss_central(e, rho, coll.D) for e in coll.monos])
    m = coll._shift(keq, u3)
    if coll.bg is not None:
        m = m - coll.bg[:, None]
    return coll.Mmono_inv @ m


def _gauss_central(e, rho, dim):
    v = rho.copy()
    for a in range(dim):
        k = e[a]
        if k % 2:
            return np.zeros_like(rho)
        v = v * (1.0 / 3.0) ** (k // 2) * (1 if k < 4 else 3)
    return v
