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
    """Rate shared by all shear moments (lbmpy ``relaxationrates.py:23-46``)."""
    rates = set()
    for moment, info in method.relaxation_info_dict.items():
        if is_shear_moment(moment, method.dim):
            rates.add(info.relaxation_rate)
    if len(rates) == 1:
        return rates.pop()
    if len(rates) > 1:
        raise ValueError(f"Shear moments are relaxed with different relaxation times: {rates}")
    all_rates = set(method.relaxation_rates)
    if len(all_rates) == 1:
        return all_rates.pop()
    raise NotImplementedError("Shear moments seem to be not relaxed separately - "
                              "Can not determine their relaxation rate automatically")


def get_bulk_relaxation_rate(method):
    for moment, info in method.relaxation_info_dict.items():
        if is_bulk_moment(moment, method.dim):
            return info.relaxation_rate
    return get_shear_relaxation_rate(method)


def relaxation_rate_scaling(omega, level_scale_factor):
    return omega / (omega / 2 + level_scale_factor * (1 - omega / 2))
