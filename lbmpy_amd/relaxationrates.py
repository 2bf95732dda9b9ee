"""Relaxation-rate helpers (lbmpy ``src/lbmpy/relaxationrates.py:6-71``)."""
from fractions import Fraction

from .moments import is_shear_moment, is_bulk_moment


def relaxation_rate_from_lattice_viscosity(nu):
    """omega = 1 / (3 nu + 1/2)."""
    return 2 / (6 * nu + 1)


def lattice_viscosity_from_relaxation_rate(omega):
    """nu = (1/omega - 1/2) / 3."""
    return (2 - omega) / (6 * omega)


def relaxation_rate_from_magic_number(hydrodynamic_relaxation_rate, magic_number=Fraction(3, 16)):
    """Second TRT rate from the magic number Lambda (default 3/16)."""
    omega = hydrodynamic_relaxation_rate
    if isinstance(omega, float):
        magic_number = float(magic_number)
    return (4 - 2 * omega) / (4 * magic_number * omega + 2 - omega)


def get_shear_relaxation_rate(method):
    """The one rate that relaxes the method's shear moments (lbmpy ``relaxationrates.py:23-46``).  A method
    without shear moments qualifies only when a single rate relaxes all of its moments.  Shear moments with
    different rates raise ValueError; no shear moment and several rates raise NotImplementedError."""
    dim = method.dim
    found = {info.relaxation_rate for moment, info in method.relaxation_info_dict.items()
             if is_shear_moment(moment, dim)}
    if len(found) > 1:
        raise ValueError(f"the shear moments of this method relax with {len(found)} different rates "
                         f"({', '.join(sorted(map(str, found)))}); a single shear rate is needed")
    if not found:
        found = set(method.relaxation_rates)
        if len(found) != 1:
            raise NotImplementedError("this method relaxes no moment as a shear moment and uses "
                                      f"{len(found)} different rates, so it has no single shear rate")
    return next(iter(found))


def get_bulk_relaxation_rate(method):
    for moment, info in method.relaxation_info_dict.items():
        if is_bulk_moment(moment, method.dim):
            return info.relaxation_rate
    return get_shear_relaxation_rate(method)


def relaxation_rate_scaling(omega, level_scale_factor):
    return omega / (omega / 2 + level_scale_factor * (1 - omega / 2))
