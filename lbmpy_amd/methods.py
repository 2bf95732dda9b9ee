"""Lattice Boltzmann method descriptions.

A method object holds what lbmpy's ``AbstractLbMethod`` subclasses hold -- the
stencil, the relaxed moments (or cumulants) with their relaxation rates and
equilibrium values, the conserved-quantity computation, the force model and
the storage format -- but in exact numeric form instead of SymPy
assignment collections.  The HIP kernels are parameterised from these
objects (``kernel_tables.py``).

Creators follow lbmpy ``src/lbmpy/methods/creationfunctions.py``:
``create_srt`` (:224-250), ``create_trt`` (:252-279), ``create_mrt_raw``
(:306-333), ``create_central_moment`` (:336-382), ``create_mrt_orthogonal``
(:464-526), ``create_cumulant`` (:530-562), ``create_with_monomial_cumulants``
(:565-584), ``create_with_default_polynomial_cumulants`` (:587-597) and
``_get_relaxation_info_dict`` (:600-706).
"""
from collections import OrderedDict, namedtuple
from fractions import Fraction

from ._poly import Poly, MOMENT_VARS, moment_poly
from .enums import CollisionSpace
from .equilibrium import (ContinuousHydrodynamicMaxwellian, DiscreteHydrodynamicMaxwellian,
                          lattice_weights, weights_by_moment_matching, EQ_VARS)
from .moments import (get_default_moment_set_for_stencil, cascaded_moment_sets_literature,
                      mrt_orthogonal_default_groups, is_even, get_order, is_shear_moment,
                      is_bulk_moment, moment_matrix, sort_moments_into_groups_of_same_order)
from .relaxationrates import relaxation_rate_from_magic_number

RelaxationInfo = namedtuple("RelaxationInfo", ["equilibrium_value", "relaxation_rate"])


class DensityVelocityComputation:
    """Density/velocity computation (lbmpy ``methods/conservedquantitycomputation.py:84-374``).

    rho = rho_0 + delta_rho where the zeroth moment of the stored PDFs is
    delta_rho for zero-centred storage and rho otherwise; u = j / rho_ref with
    rho_ref = rho (compressible) or rho_0 = 1 (incompressible), plus the force
    model's equilibrium velocity shift F/(2 rho_ref) where the model has one."""

    def __init__(self, stencil, compressible, zero_centered, force_model=None):
        self.stencil = stencil
        self.compressible = compressible
        self.zero_centered_pdfs = zero_centered
        self.force_model = force_model
        self.background_density = 1

    def density_velocity(self, pdfs):
        """(rho, u_eq, u_macroscopic) for one cell's stored PDFs (plain Python numbers)."""
        q = self.stencil.Q
        s = sum(pdfs)
        rho = 1 + s if self.zero_centered_pdfs else s
        j = [sum(pdfs[i] * self.stencil[i][a] for i in range(q)) for a in range(self.stencil.D)]
        ref = rho if self.compressible else 1
        u = [ja / ref for ja in j]
        u_eq = list(u)
        u_out = list(u)
        if self.force_model is not None:
            F = self.force_model.force
            if self.force_model.equilibrium_velocity_shift:
                u_eq = [ua + fa / (2 * ref) for ua, fa in zip(u, F)]
            u_out = [ua + fa / (2 * ref) for ua, fa in zip(u, F)]
        return rho, u_eq, u_out


class AbstractLbMethod:
    kind = None

    def __init__(self, stencil, equilibrium, relaxation_dict, cqc, force_model, zero_centered,
                 collision_space):
        self.stencil = stencil
        self.equilibrium_distribution = equilibrium
        self._relaxation_dict = OrderedDict(relaxation_dict)
        self.conserved_quantity_computation = cqc
        self.force_model = force_model
        self.zero_centered_pdfs = zero_centered
        self.collision_space = collision_space
        self._weights = None

    @property
    def dim(self):
        return self.stencil.D

    @property
    def compressible(self):
        return self.equilibrium_distribution.compressible

    @property
    def relaxation_rates(self):
        return tuple(self._relaxation_dict.values())

    @property
    def relaxation_rate_dict(self):
        return self._relaxation_dict

    @property
    def weights(self):
        if self._weights is None:
            self._weights = self._compute_weights()
        return self._weights

    def override_weights(self, weights):
        assert len(weights) == self.stencil.Q
        self._weights = [Fraction(w) for w in weights]

    def _basis(self):
        return list(self._relaxation_dict.keys())

    def _compute_weights(self):
        # moment matching against the rest state (lbmpy momentbasedmethod.py:291-307); for the
        # discrete equilibrium the background's populations are the hard-coded lattice weights
        eq = self.equilibrium_distribution
        if isinstance(eq, DiscreteHydrodynamicMaxwellian):
            return lattice_weights(self.stencil)
        bg = ContinuousHydrodynamicMaxwellian(dim=self.dim, compressible=True, deviation_only=False,
                                              order=eq.order, c_s_sq=eq.c_s_sq)
        return weights_by_moment_matching(self.stencil, self._basis(), bg)

    @property
    def moment_matrix(self):
        return moment_matrix(self._basis(), self.stencil)

    def relaxation_rate_symbols(self):
        """Free (symbolic) relaxation rates, bound at kernel call time."""
        names = []
        for r in self.relaxation_rates:
            for s in getattr(r, "free_symbols", ()):
                if s.name not in names:
                    names.append(s.name)
        return names

    def __repr__(self):
        return (f"{type(self).__name__}({self.stencil.name}, compressible={self.compressible}, "
                f"zero_centered={self.zero_centered_pdfs}, force_model={self.force_model})")


class MomentBasedLbMethod(AbstractLbMethod):
    """SRT/TRT/MRT: relax polynomial raw moments (in population or raw-moment space;
    lbmpy ``methods/momentbased/momentbasedmethod.py:16-416``)."""
    kind = "moment"

    @property
    def moments(self):
        return tuple(self._relaxation_dict.keys())

    @property
    def moment_equilibrium_values(self):
        return tuple(self.equilibrium_distribution.moment(m) for m in self.moments)

    @property
    def relaxation_info_dict(self):
        return OrderedDict((m, RelaxationInfo(v, r)) for (m, r), v in
                           zip(self._relaxation_dict.items(), self.moment_equilibrium_values))

    def set_zeroth_moment_relaxation_rate(self, rr):
        self._relaxation_dict[moment_poly((0, 0, 0))] = rr

    def set_first_moment_relaxation_rate(self, rr):
        for d in range(self.dim):
            e = [0, 0, 0]
            e[d] = 1
            m = moment_poly(tuple(e))
            assert m in self._relaxation_dict, "First moments are not relaxed separately by this method"
            self._relaxation_dict[m] = rr

    def set_conserved_moments_relaxation_rate(self, rr):
        self.set_zeroth_moment_relaxation_rate(rr)
        self.set_first_moment_relaxation_rate(rr)


class CentralMomentBasedLbMethod(AbstractLbMethod):
    """Relax polynomial central moments (lbmpy
    ``methods/momentbased/centralmomentbasedmethod.py:31-394``)."""
    kind = "central_moment"

    @property
    def moments(self):
        return tuple(self._relaxation_dict.keys())

    @property
    def moment_equilibrium_values(self):
        return tuple(self.equilibrium_distribution.central_moment(m) for m in self.moments)

    @property
    def relaxation_info_dict(self):
        return OrderedDict((m, RelaxationInfo(v, r)) for (m, r), v in
                           zip(self._relaxation_dict.items(), self.moment_equilibrium_values))


class CumulantBasedLbMethod(AbstractLbMethod):
    """Relax polynomial cumulants (lbmpy ``methods/cumulantbased/cumulantbasedmethod.py:21-402``)."""
    kind = "cumulant"

    @property
    def cumulants(self):
        return tuple(self._relaxation_dict.keys())

    moments = cumulants

    @property
    def cumulant_equilibrium_values(self):
        return tuple(self.equilibrium_distribution.cumulant(c) for c in self.cumulants)

    @property
    def relaxation_info_dict(self):
        return OrderedDict((m, RelaxationInfo(v, r)) for (m, r), v in
                           zip(self._relaxation_dict.items(), self.cumulant_equilibrium_values))


# ------------------------------------------------------------------------------ creators

def _equilibrium(stencil, continuous, compressible, deviation_only, order, c_s_sq):
    if continuous:
        return ContinuousHydrodynamicMaxwellian(dim=stencil.D, compressible=compressible,
                                                deviation_only=deviation_only, order=order, c_s_sq=c_s_sq)
    return DiscreteHydrodynamicMaxwellian(stencil, compressible=compressible, deviation_only=deviation_only,
                                          order=order, c_s_sq=c_s_sq)


def create_from_equilibrium(stencil, equilibrium, cqc, rr_dict, collision_space, zero_centered, force_model):
    if len(rr_dict) != stencil.Q:
        raise AssertionError("The number of moments has to be equal to the number of stencil entries")
    if collision_space in (CollisionSpace.POPULATIONS, CollisionSpace.RAW_MOMENTS):
        if collision_space == CollisionSpace.POPULATIONS and zero_centered and not equilibrium.deviation_only:
            raise Exception("Can only derive population-space equations for zero-centered storage"
                            " if delta equilibrium is used.")
        cls = MomentBasedLbMethod
    elif collision_space == CollisionSpace.CENTRAL_MOMENTS:
        cls = CentralMomentBasedLbMethod
    else:
        cls = CumulantBasedLbMethod
        if force_model is not None and not force_model.has_symmetric_central_moment_forcing:
            raise ValueError(f"Force model {force_model} does not offer symmetric central moment forcing.")
    return cls(stencil, equilibrium, rr_dict, cqc, force_model, zero_centered, collision_space)


def _create_moment_method(stencil, rr_dict, collision_space, continuous_equilibrium=True, compressible=False,
                          zero_centered=False, delta_equilibrium=False, force_model=None, equilibrium_order=2,
                          c_s_sq=Fraction(1, 3)):
    cqc = DensityVelocityComputation(stencil, compressible, zero_centered, force_model)
    eq = _equilibrium(stencil, continuous_equilibrium, compressible, delta_equilibrium, equilibrium_order, c_s_sq)
    return create_from_equilibrium(stencil, eq, cqc, rr_dict, collision_space, zero_centered, force_model)


def create_srt(stencil, relaxation_rate, collision_space=CollisionSpace.POPULATIONS, **kwargs):
    moments = get_default_moment_set_for_stencil(stencil)
    rr = OrderedDict((m, relaxation_rate) for m in moments)
    return _create_moment_method(stencil, rr, collision_space, **kwargs)


def create_trt(stencil, relaxation_rate_even_moments, relaxation_rate_odd_moments,
               collision_space=CollisionSpace.POPULATIONS, **kwargs):
    moments = get_default_moment_set_for_stencil(stencil)
    rr = OrderedDict((m, relaxation_rate_even_moments if is_even(m) else relaxation_rate_odd_moments)
                     for m in moments)
    return _create_moment_method(stencil, rr, collision_space, **kwargs)


def create_trt_with_magic_number(stencil, relaxation_rate, magic_number=Fraction(3, 16), **kwargs):
    return create_trt(stencil, relaxation_rate, relaxation_rate_from_magic_number(relaxation_rate, magic_number),
                      **kwargs)


def create_mrt_raw(stencil, relaxation_rates, conserved_moments=True,
                   collision_space=CollisionSpace.RAW_MOMENTS, **kwargs):
    moments = get_default_moment_set_for_stencil(stencil)
    nested = [(m,) for m in moments]
    rr = _get_relaxation_info_dict(relaxation_rates, nested, stencil.D, conserved_moments)
    return _create_moment_method(stencil, rr, collision_space, **kwargs)


def create_mrt_orthogonal(stencil, relaxation_rates, weighted=True, nested_moments=None, conserved_moments=True,
                          collision_space=CollisionSpace.RAW_MOMENTS, **kwargs):
    if not nested_moments:
        weights = lattice_weights(stencil) if weighted else None
        nested_moments = mrt_orthogonal_default_groups(stencil, weights)
    rr = _get_relaxation_info_dict(relaxation_rates, nested_moments, stencil.D, conserved_moments)
    return _create_moment_method(stencil, rr, collision_space, **kwargs)


def _split_shear_bulk(nested, dim):
    second = nested[2]
    bulk = [m for m in second if is_bulk_moment(m, dim)]
    shear = [m for m in second if is_shear_moment(m, dim)]
    assert len(shear) + len(bulk) == len(second)
    nested[2] = shear
    nested.insert(3, bulk)
    return nested


def create_central_moment(stencil, relaxation_rates, nested_moments=None, conserved_moments=True, **kwargs):
    if nested_moments and not isinstance(nested_moments[0], (list, tuple)):
        nested_moments = _split_shear_bulk(list(sort_moments_into_groups_of_same_order(nested_moments).values()),
                                           stencil.D)
    if not nested_moments:
        nested_moments = cascaded_moment_sets_literature(stencil)
    rr = _get_relaxation_info_dict(relaxation_rates, nested_moments, stencil.D, conserved_moments)
    return _create_moment_method(stencil, rr, CollisionSpace.CENTRAL_MOMENTS, **kwargs)


def create_cumulant(stencil, relaxation_rates, cumulant_groups, conserved_moments=True, zero_centered=True,
                    force_model=None, c_s_sq=Fraction(1, 3)):
    rr = _get_relaxation_info_dict(relaxation_rates, cumulant_groups, stencil.D, conserved_moments)
    cqc = DensityVelocityComputation(stencil, True, zero_centered, force_model)
    eq = ContinuousHydrodynamicMaxwellian(dim=stencil.D, compressible=True, deviation_only=False, order=2,
                                          c_s_sq=c_s_sq)
    return create_from_equilibrium(stencil, eq, cqc, rr, CollisionSpace.CUMULANTS, zero_centered, force_model)


def create_with_monomial_cumulants(stencil, relaxation_rates, conserved_moments=True, **kwargs):
    cumulants = get_default_moment_set_for_stencil(stencil)
    return create_cumulant(stencil, relaxation_rates, [(c,) for c in cumulants], conserved_moments, **kwargs)


def create_with_default_polynomial_cumulants(stencil, relaxation_rates, conserved_moments=True, **kwargs):
    return create_cumulant(stencil, relaxation_rates, cascaded_moment_sets_literature(stencil),
                           conserved_moments, **kwargs)


def _get_relaxation_info_dict(relaxation_rates, nested_moments, dim, conserved_moments=True):
    """Moment -> relaxation rate under lbmpy's rules for a rate list (methods/creationfunctions.py:600-706):

    * a callable maps each group to the rates of its moments;
    * one rate per moment: taken in order;
    * a single rate: the shear moments take it, moments of order <= 1 take 0 (or the rate, when they are not
      conserved) and every other moment relaxes with 1;
    * any other count: the shear rate (when the set has shear moments), the bulk rate (when it has a bulk
      moment), then one rate per further group that holds a moment of neither kind.  A rate is taken from the
      list when a group begins and no rate is waiting, and it waits for the next group when this one has no
      moment that needs it -- so a list that runs out raises ValueError even at such a group."""
    moments = [m for g in nested_moments for m in g]
    if callable(relaxation_rates):
        return OrderedDict((m, w) for g in nested_moments for m, w in zip(g, relaxation_rates(g)))
    rates = list(relaxation_rates)
    if len(rates) == len(moments):
        return OrderedDict(zip(moments, rates))

    def shear(m):
        return is_shear_moment(m, dim)

    def low(m):
        return get_order(m) <= 1

    if len(rates) == 1:
        w = rates[0]
        return OrderedDict((m, w if shear(m) else (0.0 if conserved_moments else w) if low(m) else 1.0)
                           for m in moments)

    supply = iter(rates)

    def take():
        for w in supply:
            return w
        raise ValueError(f"{len(rates)} relaxation rates do not cover {len(moments)} moments: pass a single "
                         "(shear) rate, one rate per moment, or the shear and bulk rates followed by one rate for "
                         "each further group of moments")

    w_shear = take() if any(shear(m) for m in moments) else None
    w_bulk = take() if any(is_bulk_moment(m, dim) for m in moments) else None
    result = OrderedDict()
    waiting = None
    for group in nested_moments:
        if waiting is None:
            waiting = take()
        spent = False
        for m in group:
            if low(m) and conserved_moments:
                result[m] = 0.0
            elif shear(m) or low(m):
                result[m] = w_shear
            elif is_bulk_moment(m, dim):
                result[m] = w_bulk
            else:
                result[m] = waiting
                spent = True
        if spent:
            waiting = None
    return result

    relaxation_rates = list(relaxation_rates)
    n = sum(len(g) for g in nested_moments)
    n_shear = sum(1 for g in nested_moments for m in g if is_shear_moment(m, dim))
    n_bulk = sum(1 for g in nested_moments for m in g if is_bulk_moment(m, dim))

    if len(relaxation_rates) == 1:
        for group in nested_moments:
            for m in group:
                if conserved_moments:
                    if get_order(m) <= 1:
                        result[m] = 0.0
                    elif is_shear_moment(m, dim):
                        result[m] = relaxation_rates[0]
                    else:
                        result[m] = 1.0
                else:
                    result[m] = relaxation_rates[0] if (is_shear_moment(m, dim) or get_order(m) <= 1) else 1.0

    if len(relaxation_rates) == n:
        it = iter(relaxation_rates)
        for group in nested_moments:
            for m in group:
                result[m] = next(it)

    if len(relaxation_rates) not in (1, n):
        try:
            it = iter(relaxation_rates)
            shear_rr = next(it) if n_shear > 0 else None
            bulk_rr = next(it) if n_bulk > 0 else None
            next_rr = True
            rr = None
            for group in nested_moments:
                if next_rr:
                    rr = next(it)
                next_rr = False
                for m in group:
                    if conserved_moments and get_order(m) <= 1:
                        result[m] = 0.0
                    elif is_shear_moment(m, dim) or (not conserved_moments and get_order(m) <= 1):
                        result[m] = shear_rr
                    elif is_bulk_moment(m, dim):
                        result[m] = bulk_rr
                    else:
                        next_rr = True
                        result[m] = rr
        except StopIteration:
            raise ValueError("Not enough relaxation rates are specified. You can either specify one relaxation "
                             "rate, which is used as the shear relaxation rate. In this case, conserved moments "
                             "are relaxed with 0, and higher-order moments are relaxed with 1. Another "
                             "possibility is to specify a relaxation rate for shear, bulk and one for each order "
                             "moment. In this case, conserved moments are also relaxed with 0. The last "
                             "possibility is to specify a relaxation rate for each moment, including conserved "
                             "moments")
    return result
