"""Method layer (host side): stencils, weights, moment bases, config normalisation.

Pins: the literal tables of the reference (stencils.py:130-266,
maxwellian_equilibrium.py:24-61, default_moment_sets.py:9-145), the properties
its tests check (tests/test_weights.py:8-34, tests/test_stencils.py,
tests/test_moments.py), and closed forms (isotropy sums, orthogonality).
"""
from fractions import Fraction

import numpy as np
import pytest

from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, ForceModel, create_lb_method, Timestep,
                       relaxation_rate_from_magic_number)
from lbmpy_amd import forcemodels
from lbmpy_amd.equilibrium import lattice_weights
from lbmpy_amd.moments import (get_default_moment_set_for_stencil, monomial_str, moment_matrix,
                               cascaded_moment_sets_literature, get_order, is_shear_moment, is_bulk_moment)
from lbmpy_amd._poly import Poly, MOMENT_VARS

STENCILS = [Stencil.D2Q9, Stencil.D3Q19, Stencil.D3Q27]


@pytest.mark.parametrize("name", STENCILS)
def test_stencil_isotropy_and_inverse(name):
    st = LBStencil(name)
    w = lattice_weights(st)
    c = np.array(st.stencil_entries)
    assert sum(w) == 1
    for a in range(st.D):
        assert sum(wi * ci[a] for wi, ci in zip(w, st)) == 0
        for b in range(st.D):
            assert sum(wi * ci[a] * ci[b] for wi, ci in zip(w, st)) == (Fraction(1, 3) if a == b else 0)
    inv = st.inverse_indices()
    for i in range(st.Q):
        assert (c[inv[i]] == -c[i]).all()
        assert inv[inv[i]] == i


def test_walberla_orderings_literal():
    # first entries of the reference's 'walberla' tables (stencils.py:133-135, 194-198, 221-227)
    assert LBStencil(Stencil.D2Q9).stencil_entries[:5] == ((0, 0), (0, 1), (0, -1), (-1, 0), (1, 0))
    d3q19 = LBStencil(Stencil.D3Q19).stencil_entries
    assert d3q19[7:11] == ((-1, 1, 0), (1, 1, 0), (-1, -1, 0), (1, -1, 0))
    d3q27 = LBStencil(Stencil.D3Q27).stencil_entries
    assert d3q27[19:] == ((1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, 1, -1), (-1, 1, -1), (1, -1, -1),
                          (-1, -1, -1))


@pytest.mark.parametrize("method", [Method.SRT, Method.TRT, Method.MRT, Method.CENTRAL_MOMENT])
@pytest.mark.parametrize("zero_centered", [False, True])
@pytest.mark.parametrize("continuous", [False, True])
@pytest.mark.parametrize("name", STENCILS)
def test_weights_match_hardcoded(method, zero_centered, continuous, name):
    """tests/test_weights.py:8-34: moment-matched weights equal the hard-coded lattice weights."""
    if method == Method.MRT and name == Stencil.D3Q27:
        pytest.skip("no default weighted-orthogonal MRT for D3Q27 (reference raises as well)")
    st = LBStencil(name)
    m = create_lb_method(LBMConfig(stencil=st, method=method, zero_centered=zero_centered,
                                   continuous_equilibrium=continuous, relaxation_rate=1.5))
    assert list(m.weights) == lattice_weights(st)


def test_hardcoded_weight_values():
    assert set(lattice_weights(LBStencil(Stencil.D2Q9))) == {Fraction(4, 9), Fraction(1, 9), Fraction(1, 36)}
    assert set(lattice_weights(LBStencil(Stencil.D3Q19))) == {Fraction(1, 3), Fraction(1, 18), Fraction(1, 36)}
    assert set(lattice_weights(LBStencil(Stencil.D3Q27))) == {Fraction(8, 27), Fraction(2, 27), Fraction(1, 54),
                                                              Fraction(1, 216)}


def test_default_moment_sets():
    d2 = [monomial_str(next(iter(p.terms))) for p in get_default_moment_set_for_stencil(LBStencil(Stencil.D2Q9))]
    # SymPy ordering key: (order, number of symbols, len(str), str)
    assert d2 == ["1", "x", "y", "x**2", "y**2", "x*y", "x**2*y", "x*y**2", "x**2*y**2"]
    d19 = get_default_moment_set_for_stencil(LBStencil(Stencil.D3Q19))
    assert len(d19) == 19
    names = {monomial_str(next(iter(p.terms))) for p in d19}
    for removed in ["x*y*z", "x*y*z**2", "x**2*y**2*z**2", "x*y**2*z**2"]:
        assert removed not in names
    m = moment_matrix(d19, LBStencil(Stencil.D3Q19))
    assert np.linalg.matrix_rank(np.array(m, dtype=float)) == 19


def test_cascaded_sets_literal_d3q27():
    groups = cascaded_moment_sets_literature(LBStencil(Stencil.D3Q27))
    assert [len(g) for g in groups] == [1, 3, 5, 1, 3, 3, 1, 2, 1, 3, 3, 1]
    x, y, z = (Poly.var(MOMENT_VARS, v) for v in MOMENT_VARS)
    assert groups[7][0] == x ** 2 * y ** 2 - 2 * x ** 2 * z ** 2 + y ** 2 * z ** 2
    assert all(is_shear_moment(p, 3) for p in groups[2])
    assert is_bulk_moment(groups[3][0], 3)


def test_mrt_d3q19_weighted_orthogonal():
    """The default D3Q19 MRT basis is weighted-orthogonal (methods/creationfunctions.py:491-516)."""
    st = LBStencil(Stencil.D3Q19)
    m = create_lb_method(LBMConfig(stencil=st, method=Method.MRT, relaxation_rate=1.8))
    M = np.array(m.moment_matrix, dtype=float)
    w = np.array([float(v) for v in m.weights])
    G = (M * w) @ M.T
    assert np.allclose(G - np.diag(np.diag(G)), 0.0)
    rates = m.relaxation_rates
    shear = [r for p, r in m.relaxation_rate_dict.items() if is_shear_moment(p, 3)]
    assert shear == [1.8] * 5
    assert sum(1 for p in m.moments if get_order(p) <= 1) == 4
    assert sorted(set(rates)) == [0.0, 1.0, 1.8]


def test_config_normalisation():
    """LBMConfig.__post_init__ rules (creationfunctions.py:347-484)."""
    c = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rate=1.8)
    assert c.relaxation_rates[1] == pytest.approx(relaxation_rate_from_magic_number(1.8))
    assert c.delta_equilibrium is True
    c = LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.8, compressible=True,
                  force=(1e-6, 0, 0))
    assert c.delta_equilibrium is False
    assert isinstance(c.force_model, forcemodels.CentralMoment)
    c = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.SRT, relaxation_rate=1.8, force=(1e-6, 0, 0))
    assert isinstance(c.force_model, forcemodels.Guo)
    c = LBMConfig(method=Method.SRT, force_model=ForceModel.LUO, force=(1e-6, 0))
    assert isinstance(c.force_model, forcemodels.Luo)
    with pytest.raises(ValueError):
        LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.8)
    with pytest.raises(ValueError):
        LBMConfig(method=Method.CUMULANT, compressible=True, delta_equilibrium=True)
    c = LBMConfig(kernel_type='aa_odd')
    assert c.streaming_pattern == 'aa' and c.timestep == Timestep.ODD
    c = LBMConfig()
    assert str(c.relaxation_rates[0]) == "omega"


def test_relaxation_dict_fallback_groups():
    """Shear, bulk, then one rate per further group (methods/creationfunctions.py:662-702)."""
    st = LBStencil(Stencil.D3Q27)
    rates = [1.9, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.75, 1.8]
    m = create_lb_method(LBMConfig(stencil=st, method=Method.CUMULANT, compressible=True, relaxation_rates=rates))
    r = list(m.relaxation_rates)
    assert r[:4] == [0.0] * 4
    assert r[4:9] == [1.9] * 5 and r[9] == 1.1
    # groups after bulk: one rate each, in order
    assert r[10:13] == [1.2] * 3 and r[13:16] == [1.3] * 3 and r[16] == 1.4
    assert r[17:19] == [1.5] * 2 and r[19] == 1.6 and r[20:23] == [1.7] * 3 and r[23:26] == [1.75] * 3
    assert r[26] == 1.8
    with pytest.raises(ValueError):
        create_lb_method(LBMConfig(stencil=st, method=Method.CUMULANT, compressible=True,
                                   relaxation_rates=rates[:5]))
