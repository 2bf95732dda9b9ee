"""Force models supported by the sweep kernels.

Restates lbmpy ``src/lbmpy/forcemodels.py``: ``Simple`` (:179-210),
``CentralMoment`` (:213-232), ``Luo`` (:235-270) and ``Guo`` (:273-316), with
the default velocity shifts of :592-597.  The population-space terms are
returned as exact coefficient tables ``(t0, t_u)`` such that the force term of
direction i is ``t0[i] + sum_a t_u[i][a] * u_a`` (all these models are affine
in the equilibrium velocity u).  Force values are plain numbers; per-cell force
fields are not part of the GPU path.
"""
from fractions import Fraction


class AbstractForceModel:
    #: shift the velocity that enters the equilibrium by F/(2 rho_ref)
    equilibrium_velocity_shift = False
    has_moment_space_forcing = True
    has_symmetric_central_moment_forcing = False

    def __init__(self, force):
        self._force = tuple(float(f) for f in force)

    @property
    def symbolic_force_vector(self):
        return self._force

    @property
    def force(self):
        return self._force

    def macroscopic_velocity_shift(self, density):
        """u_out = j/rho_ref + F/(2 rho_ref) (lbmpy ``forcemodels.py:126-133,592-593``)."""
        return [f / (2 * density) for f in self._force]

    def __repr__(self):
        return f"{type(self).__name__}({self._force})"

    def population_terms(self, method):
        raise NotImplementedError(f"{type(self).__name__} has no population-space forcing term")


def _luo_tables(method, force):
    """Luo term w_i [F.(c_i - u)/cs2 + (c_i.F)(c_i.u)/cs2^2] as affine coefficients in u."""
    cs2 = Fraction(1, 3)
    D = method.dim
    F = [Fraction(f) for f in force]
    t0, tu = [], []
    for w, c in zip(method.weights, method.stencil):
        cf = sum((F[a] * c[a] for a in range(D)), Fraction(0))
        t0.append(w * cf / cs2)
        tu.append([w * (-F[a] / cs2 + cf * c[a] / cs2 ** 2) for a in range(D)])
    return t0, tu


class Simple(AbstractForceModel):
    """(w_i/cs2) c_i.F, no equilibrium shift."""

    def population_terms(self, method):
        cs2 = Fraction(1, 3)
        F = [Fraction(f) for f in self._force]
        t0 = [w / cs2 * sum((F[a] * c[a] for a in range(method.dim)), Fraction(0))
              for w, c in zip(method.weights, method.stencil)]
        return t0, [[Fraction(0)] * method.dim for _ in t0]


class Luo(AbstractForceModel):
    def population_terms(self, method):
        return _luo_tables(method, self._force)


class Guo(AbstractForceModel):
    """Luo term times (I - D/2) in moment space; shifts the equilibrium velocity."""
    equilibrium_velocity_shift = True

    def population_terms(self, method):
        # the (I - D/2) moment-space correction is applied where the relaxation
        # rates are known (kernel table construction); here only the Luo part
        return _luo_tables(method, self._force)


class CentralMoment(AbstractForceModel):
    """Only shifts velocities; F/2 is added to the first central moments before and after
    the collision (cumulant methods only)."""
    equilibrium_velocity_shift = True
    has_moment_space_forcing = False
    has_symmetric_central_moment_forcing = True

    def population_terms(self, method):
        raise ValueError("This force model can only be combined with the Cumulant collision model")


class _Unsupported(AbstractForceModel):
    def population_terms(self, method):
        raise NotImplementedError(f"force model {type(self).__name__} is not supported by the GPU sweep")


class He(_Unsupported):
    equilibrium_velocity_shift = True


class Buick(_Unsupported):
    pass


class EDM(_Unsupported):
    pass


class ShanChen(_Unsupported):
    equilibrium_velocity_shift = True


FORCE_MODEL_BY_NAME = {
    'simple': Simple, 'luo': Luo, 'guo': Guo, 'schiller': Guo, 'buick': Buick, 'silva': Buick,
    'edm': EDM, 'kupershtokh': EDM, 'he': He, 'shanchen': ShanChen, 'centralmoment': CentralMoment,
}
