"""Kernel-creation API: ``LBMConfig``, ``LBMOptimisation``, ``create_lb_*``.

Same names, fields, defaults and normalisation rules as lbmpy
``src/lbmpy/creationfunctions.py``:

* ``LBMConfig`` (:100-484) with ``__post_init__`` (:347-484): single rate ->
  list (TRT second rate by magic number 3/16), cumulant requires
  ``compressible=True``, ``delta_equilibrium`` inference, collision-space
  inference, legacy ``kernel_type`` strings, default force model
  (CentralMoment for cumulant, Guo otherwise), force-model names.
* ``LBMOptimisation`` (:487-540).
* ``create_lb_function`` (:543-559) returns a callable sweep with ``.method``,
  ``.update_rule`` and ``.ast`` attributes; here the callable launches the
  MI355X HIP kernel through the C-ABI library (``sweep.py``).
* ``create_lb_method`` (:773-868) and ``update_with_default_parameters``
  (:920-971).

``create_lb_update_rule``/``create_lb_collision_rule``/``create_lb_ast``
return descriptive objects (the GPU path has no symbolic assignment lists);
they carry the method and the kernel specification.
"""
from dataclasses import dataclass, field, replace
from fractions import Fraction
from typing import Any, Iterable, Tuple
import warnings

from .enums import Stencil, Method, CollisionSpace, ForceModel, SubgridScaleModel, Target
from .stencils import LBStencil
from . import forcemodels
from .relaxationrates import relaxation_rate_from_magic_number
from .methods import (create_srt, create_trt, create_mrt_orthogonal, create_mrt_raw, create_central_moment,
                      create_cumulant, create_with_default_polynomial_cumulants, create_with_monomial_cumulants)

try:
    import sympy as _sp
except ImportError:  # pragma: no cover
    _sp = None


def _symbol(name):
    if _sp is None:
        raise RuntimeError("symbolic relaxation rates need sympy")
    return _sp.Symbol(name)


class Timestep(int):
    """EVEN/ODD/BOTH (lbmpy ``advanced_streaming/utility.py:18-37``)."""
    _names = {0: "Even", 1: "Odd", 2: "Both"}

    def next(self):
        return self if self == 2 else Timestep((self + 1) % 2)

    @property
    def idx(self):
        return self % 2

    def __str__(self):
        return self._names[int(self)]

    def __repr__(self):
        return f"Timestep.{self._names[int(self)].upper()}"


Timestep.EVEN = Timestep(0)
Timestep.ODD = Timestep(1)
Timestep.BOTH = Timestep(2)


@dataclass
class CollisionSpaceInfo:
    collision_space: CollisionSpace
    raw_moment_transform_class: Any = None
    central_moment_transform_class: Any = None
    cumulant_transform_class: Any = None


@dataclass
class CreateKernelConfig:
    """The subset of pystencils' ``CreateKernelConfig`` the sweep honours.

    ``data_type`` is the stored PDF type ('float64', 'float32' or 'float16');
    ``compute_type`` (an extension, default = ``data_type`` widened to at
    least float32) is the arithmetic type.  ``ghost_layers`` defaults to 1 as
    in lbmpy (``creationfunctions.py:576``); 0 is accepted for periodic or
    wall-bounded domains that need no ghost cells.

    ``stochastic_rounding`` (an extension, FP16 storage only) rounds every stored population to one of its two
    neighbouring FP16 values with probabilities that make the rounding unbiased, from a counter-based hash of
    (cell, direction, launch) seeded by ``rounding_seed``: runs are reproducible for a seed, and increments
    smaller than half an FP16 ulp (a weak body force) survive on average instead of being rounded away."""
    target: Target = Target.CPU
    data_type: str = "float64"
    compute_type: str = None
    ghost_layers: int = 1
    cpu_openmp: Any = False
    gpu_indexing_params: Any = None
    backend: Any = None
    cpu_vectorize_info: Any = None
    stochastic_rounding: bool = False
    rounding_seed: int = 0

    def __post_init__(self):
        if isinstance(self.data_type, dict):
            self.data_type = next(iter(self.data_type.values()))
        self.data_type = _normalise_dtype(self.data_type)
        if self.compute_type is None:
            self.compute_type = "float64" if self.data_type == "float64" else "float32"
        self.compute_type = _normalise_dtype(self.compute_type)
        if self.stochastic_rounding and self.data_type != "float16":
            raise ValueError("stochastic rounding applies to FP16 storage (data_type='float16')")


def _normalise_dtype(dt):
    s = str(dt).replace("torch.", "").replace("numpy.", "")
    aliases = {"double": "float64", "float": "float32", "half": "float16", "fp64": "float64", "fp32": "float32",
               "fp16": "float16", "f64": "float64", "f32": "float32", "f16": "float16"}
    s = aliases.get(s, s)
    if s not in ("float64", "float32", "float16"):
        raise ValueError(f"unsupported PDF data type {dt!r}")
    return s


@dataclass
class LBMConfig:
    stencil: Any = LBStencil(Stencil.D2Q9)
    method: Method = Method.SRT
    relaxation_rates: Iterable = None
    relaxation_rate: Any = None
    compressible: bool = False
    zero_centered: bool = True
    delta_equilibrium: bool = None
    equilibrium_order: int = 2
    c_s_sq: Any = Fraction(1, 3)
    weighted: bool = True
    nested_moments: Any = None
    force_model: Any = None
    force: Tuple = (0, 0, 0)
    continuous_equilibrium: bool = True
    maxwellian_moments: bool = None
    initial_velocity: Tuple = None
    galilean_correction: bool = False
    fourth_order_correction: Any = False
    collision_space_info: Any = None
    entropic: bool = False
    entropic_newton_iterations: int = None
    omega_output_field: Any = None
    eddy_viscosity_field: Any = None
    subgrid_scale_model: Any = None
    cassons: Any = False
    fluctuating: dict = False
    temperature: Any = None
    psm_config: Any = None
    output: dict = field(default_factory=dict)
    velocity_input: Any = None
    density_input: Any = None
    conserved_moments: bool = True
    kernel_type: Any = 'default_stream_collide'
    streaming_pattern: str = 'pull'
    timestep: Any = Timestep.BOTH
    field_name: str = 'src'
    temporary_field_name: str = 'dst'
    lb_method: Any = None
    collision_rule: Any = None
    update_rule: Any = None
    ast: Any = None

    def __post_init__(self):
        if isinstance(self.method, str):
            new_method = Method[self.method.upper()]
            warnings.warn(f'Method "{self.method}" as str is deprecated. Use {new_method} instead')
            self.method = new_method
        if self.maxwellian_moments is not None:
            warnings.warn("Argument 'maxwellian_moments' is deprecated. Use `continuous_equilibrium` instead.")
            self.continuous_equilibrium = self.maxwellian_moments
        if not isinstance(self.stencil, LBStencil):
            self.stencil = LBStencil(self.stencil)

        if self.relaxation_rates is None and self.relaxation_rate is None:
            self.relaxation_rate = _symbol("omega")
        if self.relaxation_rate is not None:
            if self.method in (Method.TRT, Method.TRT_KBC_N1, Method.TRT_KBC_N2, Method.TRT_KBC_N3,
                               Method.TRT_KBC_N4):
                self.relaxation_rates = [self.relaxation_rate,
                                         relaxation_rate_from_magic_number(self.relaxation_rate)]
            else:
                self.relaxation_rates = [self.relaxation_rate]

        if not self.compressible and self.method in (Method.MONOMIAL_CUMULANT, Method.CUMULANT):
            raise ValueError("Incompressible cumulant-based methods are not supported (yet).")
        if self.zero_centered and self.entropic:
            raise ValueError("Entropic methods can only be created with `zero_centered=False`.")

        if self.delta_equilibrium is not None:
            if self.delta_equilibrium:
                if not self.zero_centered:
                    raise ValueError("`delta_equilibrium=True` requires `zero_centered=True`!")
                if self.method in (Method.MONOMIAL_CUMULANT, Method.CUMULANT):
                    raise ValueError("Cannot create a cumulant-based method from a delta-equilibrium!")
        else:
            if self.zero_centered:
                self.delta_equilibrium = self.method not in (Method.CENTRAL_MOMENT, Method.MONOMIAL_CUMULANT,
                                                             Method.CUMULANT)
            else:
                self.delta_equilibrium = False

        if isinstance(self.collision_space_info, CollisionSpace):
            self.collision_space_info = CollisionSpaceInfo(self.collision_space_info)
        if self.collision_space_info is not None:
            if (self.entropic or self.fluctuating) \
                    and self.collision_space_info.collision_space != CollisionSpace.POPULATIONS:
                raise ValueError("Entropic and Fluctuating methods are only available in population space.")
            elif not self.collision_space_info.collision_space.compatible(self.method):
                raise ValueError("Given method is not compatible with given collision space.")
        else:
            if self.method in {Method.SRT, Method.TRT, Method.TRT_KBC_N1, Method.TRT_KBC_N2, Method.TRT_KBC_N3,
                               Method.TRT_KBC_N4} or self.entropic or self.fluctuating:
                cs = CollisionSpace.POPULATIONS
            elif self.method in {Method.MRT_RAW, Method.MRT}:
                cs = CollisionSpace.RAW_MOMENTS
            elif self.method == Method.CENTRAL_MOMENT:
                cs = CollisionSpace.CENTRAL_MOMENTS
            elif self.method in {Method.MONOMIAL_CUMULANT, Method.CUMULANT}:
                cs = CollisionSpace.CUMULANTS
            else:
                raise Exception(f"No default collision space is given for method {self.method}.")
            self.collision_space_info = CollisionSpaceInfo(cs)

        legacy = {'stream_pull_collide': ('pull', Timestep.BOTH), 'collide_stream_push': ('push', Timestep.BOTH),
                  'aa_even': ('aa', Timestep.EVEN), 'aa_odd': ('aa', Timestep.ODD),
                  'esotwist_even': ('esotwist', Timestep.EVEN), 'esotwist_odd': ('esotwist', Timestep.ODD)}
        if isinstance(self.kernel_type, str) and self.kernel_type in legacy:
            self.streaming_pattern, self.timestep = legacy[self.kernel_type]
            self.kernel_type = 'default_stream_collide'
        if not isinstance(self.timestep, Timestep):
            self.timestep = Timestep(int(self.timestep))

        force_not_zero = any(f != 0 for f in self.force)
        if self.force_model is None and force_not_zero:
            if self.method == Method.CUMULANT:
                self.force_model = forcemodels.CentralMoment(self.force[:self.stencil.D])
            else:
                self.force_model = forcemodels.Guo(self.force[:self.stencil.D])
        if isinstance(self.force_model, str):
            new_fm = ForceModel[self.force_model.upper()]
            warnings.warn(f'ForceModel "{self.force_model}" as str is deprecated. Use {new_fm} instead',
                          DeprecationWarning)
            self.force_model = forcemodels.FORCE_MODEL_BY_NAME[new_fm.name.lower()](self.force[:self.stencil.D])
        elif isinstance(self.force_model, ForceModel):
            cls = forcemodels.FORCE_MODEL_BY_NAME[self.force_model.name.lower()]
            self.force_model = cls(self.force[:self.stencil.D])
        elif isinstance(self.force_model, type) and issubclass(self.force_model, forcemodels.AbstractForceModel):
            self.force_model = self.force_model(self.force[:self.stencil.D])

        if self.density_input is not None or self.velocity_input is not None:
            self.conserved_moments = False


@dataclass
class LBMOptimisation:
    cse_pdfs: bool = False
    cse_global: bool = False
    simplification: Any = 'auto'
    pre_simplification: bool = True
    split: bool = False
    field_size: Any = None
    field_layout: str = 'fzyx'
    symbolic_field: Any = None
    symbolic_temporary_field: Any = None
    builtin_periodicity: Tuple = (False, False, False)


def update_with_default_parameters(params, opt_params=None, lbm_config=None, lbm_optimisation=None, config=None):
    """Map legacy keyword arguments onto the three config objects (lbmpy ``creationfunctions.py:920-971``)."""
    ps_params = ['target', 'backend', 'cpu_openmp', 'double_precision', 'gpu_indexing', 'gpu_indexing_params',
                 'cpu_vectorize_info', 'data_type', 'compute_type', 'ghost_layers']
    config_params = {k: v for k, v in (opt_params or {}).items() if k in ps_params}
    if 'double_precision' in config_params:
        config_params['data_type'] = 'float64' if config_params.pop('double_precision') else 'float32'
    config_params.pop('gpu_indexing', None)
    if not config:
        config = CreateKernelConfig(**config_params)
    else:
        for k, v in config_params.items():
            if not hasattr(config, k):
                raise KeyError(f'{v} is not a valid kwarg. Please look in CreateKernelConfig for valid settings')
        config = replace(config, **config_params)

    lbm_opt_keys = ['cse_pdfs', 'cse_global', 'simplification', 'pre_simplification', 'split', 'field_size',
                    'field_layout', 'symbolic_field', 'symbolic_temporary_field', 'builtin_periodicity']
    opt_dict = {k: v for k, v in (opt_params or {}).items() if k in lbm_opt_keys}
    if not lbm_optimisation:
        lbm_optimisation = LBMOptimisation(**opt_dict)
    else:
        for k, v in opt_dict.items():
            if not hasattr(lbm_optimisation, k):
                raise KeyError(f'{v} is not a valid kwarg. Please look in LBMOptimisation for valid settings')
        lbm_optimisation = replace(lbm_optimisation, **opt_dict)

    params = params or {}
    if not lbm_config:
        lbm_config = LBMConfig(**params)
    else:
        for k, v in params.items():
            if not hasattr(lbm_config, k):
                raise KeyError(f'{v} is not a valid kwarg. Please look in LBMConfig for valid settings')
        lbm_config = replace(lbm_config, **params)
    return lbm_config, lbm_optimisation, config


def create_lb_method(lbm_config=None, **params):
    """Create the method description (lbmpy ``creationfunctions.py:773-868``)."""
    lbm_config, _, _ = update_with_default_parameters(params, lbm_config=lbm_config)
    rates = lbm_config.relaxation_rates
    common = dict(compressible=lbm_config.compressible, zero_centered=lbm_config.zero_centered,
                  delta_equilibrium=lbm_config.delta_equilibrium, equilibrium_order=lbm_config.equilibrium_order,
                  force_model=lbm_config.force_model, continuous_equilibrium=lbm_config.continuous_equilibrium,
                  c_s_sq=Fraction(lbm_config.c_s_sq))
    cspace = lbm_config.collision_space_info.collision_space
    m = lbm_config.method
    if lbm_config.psm_config is not None or lbm_config.fluctuating or lbm_config.entropic or lbm_config.cassons:
        raise NotImplementedError("PSM, fluctuating, entropic and Cassons methods are not part of the GPU sweep")
    if m == Method.SRT:
        assert len(rates) >= 1, "Not enough relaxation rates"
        method = create_srt(lbm_config.stencil, rates[0], collision_space=cspace, **common)
    elif m == Method.TRT:
        assert len(rates) >= 2, "Not enough relaxation rates"
        method = create_trt(lbm_config.stencil, rates[0], rates[1], collision_space=cspace, **common)
    elif m == Method.MRT:
        method = create_mrt_orthogonal(lbm_config.stencil, rates, weighted=lbm_config.weighted,
                                       nested_moments=lbm_config.nested_moments,
                                       conserved_moments=lbm_config.conserved_moments, collision_space=cspace,
                                       **common)
    elif m == Method.MRT_RAW:
        method = create_mrt_raw(lbm_config.stencil, rates, conserved_moments=lbm_config.conserved_moments,
                                collision_space=cspace, **common)
    elif m == Method.CENTRAL_MOMENT:
        method = create_central_moment(lbm_config.stencil, rates, nested_moments=lbm_config.nested_moments,
                                       conserved_moments=lbm_config.conserved_moments, **common)
    elif m == Method.CUMULANT:
        if lbm_config.fourth_order_correction:
            # lbmpy's own condition (creationfunctions.py:835-846); a non-D3Q27 stencil that passes it
            # fails at the collision rule, as in lbmpy
            if lbm_config.stencil.D != 3 and lbm_config.stencil.Q != 27:
                raise ValueError("Fourth-order correction can only be applied to D3Q27 cumulant methods.")
            assert len(rates) <= 2, "Optimal parametrisation for fourth-order cumulants needs either one " \
                                    "or two relaxation rates, associated with the shear (and bulk) " \
                                    "viscosity. All other relaxation rates are automatically chosen optimally"
            # the method carries symbolic rates; their values are assigned with the collision rule
            from .kernel_tables import CORRECTION_RATE_NAMES
            rates = [_symbol(n) for n in CORRECTION_RATE_NAMES]
        kw = dict(zero_centered=lbm_config.zero_centered, force_model=lbm_config.force_model,
                  c_s_sq=Fraction(lbm_config.c_s_sq))
        if lbm_config.nested_moments is not None:
            method = create_cumulant(lbm_config.stencil, rates, lbm_config.nested_moments,
                                     conserved_moments=lbm_config.conserved_moments, **kw)
        else:
            method = create_with_default_polynomial_cumulants(lbm_config.stencil, rates, **kw)
    elif m == Method.MONOMIAL_CUMULANT:
        method = create_with_monomial_cumulants(lbm_config.stencil, rates,
                                                conserved_moments=lbm_config.conserved_moments,
                                                zero_centered=lbm_config.zero_centered,
                                                force_model=lbm_config.force_model)
    else:
        raise ValueError("Failed to create LB method. Please use lbmpy.enums.Method for the creation")
    lbm_config.lb_method = method
    return method


class CollisionRule:
    """Stand-in for lbmpy's ``LbmCollisionRule``: the method plus the relaxation info."""

    def __init__(self, method, lbm_config):
        self.method = method
        self.lbm_config = lbm_config
        self.simplification_hints = {}

    def __repr__(self):
        return f"CollisionRule({self.method!r})"


def create_lb_collision_rule(lb_method=None, lbm_config=None, lbm_optimisation=None, config=None,
                             optimization=None, **kwargs):
    lbm_config, lbm_optimisation, config = update_with_default_parameters(kwargs, optimization, lbm_config,
                                                                          lbm_optimisation, config)
    if lbm_config.lb_method is not None:
        lb_method = lbm_config.lb_method
    if lb_method is None:
        lb_method = create_lb_method(lbm_config)
    _check_cumulant_corrections(lb_method, lbm_config)
    rule = CollisionRule(lb_method, lbm_config)
    lbm_config.collision_rule = rule
    return rule


def _check_cumulant_corrections(method, lbm_config):
    """The checks lbmpy makes when it adds the cumulant corrections to a collision rule
    (creationfunctions.py:692-707, galilean_correction.py:21-37, fourth_order_correction.py:21-36); the
    fourth-order correction gets a bulk rate of 1 when only the shear rate is given."""
    from .kernel_tables import cumulant_corrections
    if lbm_config.galilean_correction:
        if method.kind != "cumulant" or method.stencil.name != "D3Q27":
            raise ValueError("Galilean correction is only defined for D3Q27 cumulant methods.")
    if lbm_config.fourth_order_correction:
        if len(lbm_config.relaxation_rates) == 1:
            lbm_config.relaxation_rates = list(lbm_config.relaxation_rates) + [1]
        if method.kind != "cumulant" or method.stencil.name != "D3Q27":
            raise ValueError("Fourth-order correction is only defined for D3Q27 cumulant methods.")
    return cumulant_corrections(lbm_config)


def create_lb_update_rule(collision_rule=None, lbm_config=None, lbm_optimisation=None, config=None,
                          optimization=None, **kwargs):
    lbm_config, lbm_optimisation, config = update_with_default_parameters(kwargs, optimization, lbm_config,
                                                                          lbm_optimisation, config)
    from .sweep import UpdateRule
    if lbm_config.collision_rule is not None:
        collision_rule = lbm_config.collision_rule
    if collision_rule is None:
        collision_rule = create_lb_collision_rule(lbm_config.lb_method, lbm_config=lbm_config,
                                                  lbm_optimisation=lbm_optimisation, config=config)
    rule = UpdateRule(collision_rule.method, lbm_config, lbm_optimisation, config)
    lbm_config.update_rule = rule
    return rule


def create_lb_ast(update_rule=None, lbm_config=None, lbm_optimisation=None, config=None, optimization=None,
                  **kwargs):
    lbm_config, lbm_optimisation, config = update_with_default_parameters(kwargs, optimization, lbm_config,
                                                                          lbm_optimisation, config)
    from .sweep import KernelAst
    if lbm_config.update_rule is not None:
        update_rule = lbm_config.update_rule
    if update_rule is None:
        update_rule = create_lb_update_rule(lbm_config.collision_rule, lbm_config=lbm_config,
                                            lbm_optimisation=lbm_optimisation, config=config)
    ast = KernelAst(update_rule)
    lbm_config.ast = ast
    return ast


def create_lb_function(ast=None, lbm_config=None, lbm_optimisation=None, config=None, optimization=None,
                       **kwargs):
    """Create the sweep callable (lbmpy ``creationfunctions.py:543-559``)."""
    lbm_config, lbm_optimisation, config = update_with_default_parameters(kwargs, optimization, lbm_config,
                                                                          lbm_optimisation, config)
    if lbm_config.ast is not None:
        ast = lbm_config.ast
    if ast is None:
        ast = create_lb_ast(lbm_config.update_rule, lbm_config=lbm_config, lbm_optimisation=lbm_optimisation,
                            config=config)
    res = ast.compile()
    res.method = ast.method
    res.update_rule = ast.update_rule
    res.ast = ast
    return res
