"""Macroscopic setter and getter kernels on the GPU (lbmpy ``macroscopic_value_kernels.py``).

lbmpy offers two forms:
- ``macroscopic_values_setter`` / ``macroscopic_values_getter`` (``pdf_initialization_assignments``,
  :29-63) return assignment collections that the caller compiles with pystencils;
- ``compile_macroscopic_values_setter`` / ``compile_macroscopic_values_getter`` (:88-242) return compiled
  functions.

Here both forms run the HIP kernels ``lbm355_set_equilibrium`` / ``lbm355_get_macroscopic`` on PyTorch ROCm
tensors.  The assignment-returning functions give an object whose ``compile()`` returns the kernel, so
``create_kernel``-free call sites read the same.  The PDFs are read or written through the write accessor of
the kernel that ran last (``previous_timestep``) of any streaming pattern: pull, push, AA, EsoTwist,
EsoPull, EsoPush (``get_field_accesses``, :14-26).

The arithmetic is FP64 whatever the storage type.  Density and velocity fields may be numbers or tensors
shaped like the PDF array without its last axis (velocity: plus a last axis of D components); they go through
FP64 staging fields in the PDF layout.  Setting pre-collision PDFs (``set_pre_collision_pdfs=True``) and
``iteration_slice`` are not available.  Neighbour accesses of in-place patterns wrap on the axes given by
``periodicity`` (default: periodic when there is no ghost layer) and reach into the ghost layer otherwise.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from .creationfunctions import Timestep
from .enums import Target
from .kernel_tables import method_table, family_of, STENCIL_CODES, bind_rate
from .relaxationrates import get_shear_relaxation_rate


def _pattern_code(streaming_pattern, previous_timestep):
    """Kernel code of the kernel that ran last (advanced_streaming/utility.py:52-68 semantics)."""
    if streaming_pattern == "pull":
        return nat.PULL
    if streaming_pattern == "push":
        return nat.PUSH
    if streaming_pattern not in nat.INPLACE_EVEN:
        raise ValueError("Invalid value of parameter 'streaming_pattern'.", streaming_pattern)
    t = int(previous_timestep)
    if t not in (0, 1):
        raise ValueError(f"Invalid timestep for streaming pattern {streaming_pattern}: {Timestep(t)}")
    return nat.INPLACE_EVEN[streaming_pattern] + t


class _Context:
    """The pieces both kernels need: the method's descriptor bits and FP64 staging fields."""

    def __init__(self, method, ghost_layers, periodicity, target):
        if target != Target.GPU:
            raise NotImplementedError("this package runs the macroscopic kernels on the GPU (Target.GPU)")
        self.method = method
        self.D = method.dim
        self.Q = method.stencil.Q
        self.g = int(ghost_layers)
        if self.g not in (0, 1):
            raise ValueError("ghost_layers must be 0 or 1")
        if periodicity is None:
            periodicity = (self.g == 0,) * self.D
        self.periodicity = tuple(bool(p) for p in periodicity)[:self.D]
        if not all(self.periodicity) and self.g == 0:
            raise ValueError("non-periodic axes need a ghost layer")
        self._staging = {}

    def check_pdfs(self, pdfs):
        if not (isinstance(pdfs, torch.Tensor) and pdfs.is_cuda):
            raise TypeError("pdfs must be a CUDA (ROCm) torch tensor indexed [x, y, (z,) q]")
        if pdfs.dim() != self.D + 1 or pdfs.shape[-1] != self.Q or pdfs.stride(0) != 1:
            raise ValueError(f"pdfs must be an fzyx tensor [x, y{', z' if self.D == 3 else ''}, q] with q = {self.Q}")
        if str(pdfs.dtype).replace("torch.", "") not in nat.DTYPE_CODES:
            raise ValueError(f"unsupported PDF dtype {pdfs.dtype}")

    def fields(self, pdfs):
        """FP64 staging density [x, y, (z)] and velocity [x, y, (z), D], x contiguous, pdf-shaped."""
        key = (tuple(pdfs.shape[:-1]), pdfs.device)
        f = self._staging.get(key)
        if f is None:
            shape = tuple(reversed(pdfs.shape[:-1]))
            rho = torch.zeros(shape, dtype=torch.float64, device=pdfs.device).permute(*reversed(range(self.D)))
            u = torch.zeros((self.D,) + shape, dtype=torch.float64, device=pdfs.device)
            u = u.permute(*reversed(range(1, self.D + 1)), 0)
            f = self._staging[key] = (rho, u)
        return f

    def strain_field(self, pdfs):
        """FP64 staging strain-rate tensor [x, y, (z), D*D], x contiguous, pdf-shaped."""
        key = ("strain", tuple(pdfs.shape[:-1]), pdfs.device)
        f = self._staging.get(key)
        if f is None:
            shape = tuple(reversed(pdfs.shape[:-1]))
            f = torch.zeros((self.D * self.D,) + shape, dtype=torch.float64, device=pdfs.device)
            f = self._staging[key] = f.permute(*reversed(range(1, self.D + 1)), 0)
        return f

    def descriptor(self, pdfs, rho, u):
        m = self.method
        D, g = self.D, self.g
        d = nat.SweepDesc()
        d.stencil = STENCIL_CODES[m.stencil.name]
        d.family = family_of(m)
        d.pattern = nat.PULL
        d.store_dtype = nat.DTYPE_CODES[str(pdfs.dtype).replace("torch.", "")]
        d.compute_dtype = nat.F64
        sizes = [pdfs.shape[a] - 2 * g for a in range(D)] + [1] * (3 - D)
        if min(sizes) <= 0:
            raise ValueError("PDF array smaller than its ghost layers")
        d.nx, d.ny, d.nz = sizes
        d.ghost_layers = g
        for a in range(3):
            d.face[a] = nat.FACE_PERIODIC if a >= D or self.periodicity[a] else nat.FACE_GHOST
        d.stride_y = pdfs.stride(1)
        d.stride_z = pdfs.stride(2) if D == 3 else 0
        d.stride_q = pdfs.stride(D)
        d.out_stride_y = rho.stride(1)
        d.out_stride_z = rho.stride(2) if D == 3 else 0
        d.out_stride_c = u.stride(D)
        flags = 0
        if m.zero_centered_pdfs:
            flags |= nat.FLAG_ZERO_CENTERED
        if m.compressible:
            flags |= nat.FLAG_COMPRESSIBLE
        fm = m.force_model
        if fm is not None:
            flags |= nat.FLAG_FORCE
            if fm.equilibrium_velocity_shift:
                flags |= nat.FLAG_EQ_SHIFT
            for a in range(D):
                d.force[a] = float(fm.force[a])
        d.flags = flags
        return d

    def table(self, device):
        t = getattr(self, "_table", None)
        if t is None:
            # the equilibrium does not depend on the relaxation rates: any binding of their symbols will do
            bind = {s: 1.0 for s in self.method.relaxation_rate_symbols()}
            arr, _ = method_table(self.method, bind, np.float64)
            t = self._table = torch.from_numpy(arr).to(device)
        return t


def _check_args(iteration_slice, set_pre_collision):
    if iteration_slice is not None:
        raise NotImplementedError("iteration_slice is not available on the GPU macroscopic kernels")
    if set_pre_collision:
        raise NotImplementedError("only post-collision PDFs (the previous kernel's write accessor) are "
                                  "available on the GPU macroscopic kernels")


def _set(ctx, pdfs, density, velocity, code):
    ctx.check_pdfs(pdfs)
    rho, u = ctx.fields(pdfs)
    rho.copy_(torch.as_tensor(density, dtype=torch.float64, device=pdfs.device).expand_as(rho)
              if not isinstance(density, torch.Tensor) or density.dim() == 0 else density)
    vel = velocity if isinstance(velocity, torch.Tensor) else torch.as_tensor(
        np.asarray(velocity, dtype=np.float64), device=pdfs.device)
    u.copy_(vel.expand_as(u) if vel.dim() == 1 else vel)
    d = ctx.descriptor(pdfs, rho, u)
    lib = nat.load()
    nat.check(lib.lbm355_set_equilibrium(ctypes.byref(d), ctypes.c_void_p(pdfs.data_ptr()),
                                         ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                                         ctypes.c_void_p(ctx.table(pdfs.device).data_ptr()), code,
                                         nat.current_stream_handle()))


def _get(ctx, pdfs, density, velocity, code, density_deviation=None):
    ctx.check_pdfs(pdfs)
    rho, u = ctx.fields(pdfs)
    d = ctx.descriptor(pdfs, rho, u)
    lib = nat.load()
    nat.check(lib.lbm355_get_macroscopic(ctypes.byref(d), ctypes.c_void_p(pdfs.data_ptr()),
                                         ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(u.data_ptr()), code,
                                         nat.current_stream_handle()))
    g = ctx.g
    inner = tuple(slice(g, s - g) for s in pdfs.shape[:-1])
    if density is not None:
        density[inner] = rho[inner].to(density.dtype)
    if density_deviation is not None:
        density_deviation[inner] = (rho[inner] - 1.0).to(density_deviation.dtype)
    if velocity is not None:
        velocity[inner] = u[inner].to(velocity.dtype)


def _strain(ctx, pdfs, out, code, bindings):
    ctx.check_pdfs(pdfs)
    if out is None:
        raise ValueError("give the strain-rate output tensor")
    rho, _ = ctx.fields(pdfs)
    S = ctx.strain_field(pdfs)
    d = ctx.descriptor(pdfs, rho, S)
    rate = get_shear_relaxation_rate(ctx.method)
    names = {str(k): float(v) for k, v in bindings.items()}
    omega = float(bind_rate(rate, names))
    lib = nat.load()
    nat.check(lib.lbm355_get_strain_rate(ctypes.byref(d), ctypes.c_void_p(pdfs.data_ptr()),
                                         ctypes.c_void_p(ctx.table(pdfs.device).data_ptr()),
                                         ctypes.c_void_p(S.data_ptr()), omega, code, nat.current_stream_handle()))
    g = ctx.g
    inner = tuple(slice(g, n - g) for n in pdfs.shape[:-1])
    out[inner] = S[inner].to(out.dtype)


def compile_macroscopic_values_setter(lb_method, quantities_to_set, pdf_arr=None, ghost_layers=1,
                                      iteration_slice=None, field_layout='numpy', target=Target.GPU,
                                      streaming_pattern='pull', previous_timestep=Timestep.BOTH, periodicity=None):
    """A function ``setter(pdfs)`` that writes the equilibrium of ``quantities_to_set`` ({'density': value or
    tensor, 'velocity': tuple or tensor}) into ``pdfs`` (lbmpy macroscopic_value_kernels.py:176-240)."""
    _check_args(iteration_slice, False)
    unknown = set(quantities_to_set) - {"density", "velocity"}
    if unknown:
        raise ValueError(f"No such conserved quantity: {sorted(unknown)}, conserved quantities are "
                         "['density', 'velocity']")
    ctx = _Context(lb_method, ghost_layers, periodicity, target)
    code = _pattern_code(streaming_pattern, previous_timestep)
    density = quantities_to_set.get("density", 1.0)
    velocity = quantities_to_set.get("velocity", (0.0,) * lb_method.dim)

    def setter(pdfs, **kwargs):
        if pdf_arr is not None and (pdfs.shape != pdf_arr.shape or pdfs.stride() != pdf_arr.stride()):
            raise AssertionError("Pdf array not matching blueprint which was used to compile")
        _set(ctx, pdfs, kwargs.get("density", density), kwargs.get("velocity", velocity), code)

    return setter


def compile_macroscopic_values_getter(lb_method, output_quantities, pdf_arr=None, ghost_layers=1,
                                      iteration_slice=None, field_layout='numpy', target=Target.GPU,
                                      streaming_pattern='pull', previous_timestep=Timestep.BOTH, periodicity=None):
    """A function ``getter(pdfs, density=..., density_deviation=..., velocity=...)`` that fills the requested
    output tensors over the interior (lbmpy macroscopic_value_kernels.py:88-155)."""
    _check_args(iteration_slice, False)
    if not isinstance(output_quantities, (list, tuple)):
        output_quantities = [output_quantities]
    unknown = [q for q in output_quantities if q not in ("density", "density_deviation", "velocity")]
    if unknown:
        raise ValueError(f"No such conserved quantity: {unknown}, conserved quantities are "
                         "['density', 'density_deviation', 'velocity']")
    ctx = _Context(lb_method, ghost_layers, periodicity, target)
    code = _pattern_code(streaming_pattern, previous_timestep)

    def getter(pdfs, **kwargs):
        if pdf_arr is not None and (pdfs.shape != pdf_arr.shape or pdfs.stride() != pdf_arr.stride()):
            raise AssertionError("Pdf array not matching blueprint which was used to compile")
        if not set(output_quantities).issubset(kwargs.keys()):
            raise ValueError(f"You have to specify the output field for each of the following quantities: "
                             f"{output_quantities}")
        _get(ctx, pdfs, kwargs.get("density"), kwargs.get("velocity"), code, kwargs.get("density_deviation"))

    return getter


class MacroscopicKernel:
    """The GPU kernel behind ``macroscopic_values_setter`` / ``macroscopic_values_getter``.  It stands where
    lbmpy returns assignments for pystencils' ``create_kernel``: ``compile()`` returns the kernel itself, and a
    call runs it with the fields given at creation, each replaceable by keyword (``pdfs``, ``density``,
    ``velocity``)."""

    def __init__(self, kind, ctx, code, pdfs, density, velocity, strain_rate_tensor=None):
        self.kind = kind
        self._ctx = ctx
        self._code = code
        self._fields = {"pdfs": pdfs, "density": density, "velocity": velocity,
                        "strain_rate_tensor": strain_rate_tensor}

    def compile(self):
        return self

    def __call__(self, **kwargs):
        fields = {k: v for k, v in kwargs.items() if k in self._fields}
        bindings = {k: v for k, v in kwargs.items() if k not in self._fields}
        f = dict(self._fields, **fields)
        if f["pdfs"] is None:
            raise ValueError("no PDF tensor given")
        if self.kind == "strain":
            _strain(self._ctx, f["pdfs"], f["strain_rate_tensor"], self._code, bindings)
        elif bindings:
            raise TypeError(f"unexpected arguments {sorted(bindings)}")
        elif self.kind == "setter":
            _set(self._ctx, f["pdfs"], f["density"], f["velocity"], self._code)
        else:
            _get(self._ctx, f["pdfs"], f["density"], f["velocity"], self._code)


def pdf_initialization_assignments(lb_method, density, velocity, pdfs, streaming_pattern='pull',
                                   previous_timestep=Timestep.BOTH, set_pre_collision_pdfs=False, ghost_layers=1,
                                   periodicity=None):
    """Equilibrium initialisation of ``pdfs`` (lbmpy macroscopic_value_kernels.py:29-45)."""
    _check_args(None, set_pre_collision_pdfs)
    ctx = _Context(lb_method, ghost_layers, periodicity, Target.GPU)
    return MacroscopicKernel("setter", ctx, _pattern_code(streaming_pattern, previous_timestep), pdfs, density,
                             velocity)


macroscopic_values_setter = pdf_initialization_assignments


def macroscopic_values_getter(lb_method, density, velocity, pdfs, streaming_pattern='pull',
                              previous_timestep=Timestep.BOTH, use_pre_collision_pdfs=False, ghost_layers=1,
                              periodicity=None):
    """Density and / or velocity of ``pdfs`` into the given tensors (lbmpy macroscopic_value_kernels.py:48-60)."""
    _check_args(None, use_pre_collision_pdfs)
    if density is None and velocity is None:
        raise ValueError("give a density or a velocity output")
    ctx = _Context(lb_method, ghost_layers, periodicity, Target.GPU)
    return MacroscopicKernel("getter", ctx, _pattern_code(streaming_pattern, previous_timestep), pdfs, density,
                             velocity)


def strain_rate_tensor_getter(lb_method, strain_rate_tensor, pdfs, streaming_pattern='pull',
                              previous_timestep=Timestep.BOTH, use_pre_collision_pdfs=False, ghost_layers=1,
                              periodicity=None):
    """The strain-rate tensor S = -1.5 (omega_s / rho_ref) Pi_neq of the populations the kernel that ran at
    ``previous_timestep`` wrote, into ``strain_rate_tensor`` [x, y, (z), D*D] (component a*D + b), over the
    interior (macroscopic_value_kernels.py:67-85).  Pi_neq is the second moment of f - f_eq(rho, u) with rho
    and the equilibrium velocity u computed from the same populations, as the collision computes them;
    rho_ref is rho for compressible methods and 1 otherwise; omega_s is the method's shear rate, bound by
    keyword at the call when it is a symbol.  Where lbmpy returns assignments, this returns a GPU kernel whose
    ``compile()`` is itself."""
    _check_args(None, use_pre_collision_pdfs)
    ctx = _Context(lb_method, ghost_layers, periodicity, Target.GPU)
    return MacroscopicKernel("strain", ctx, _pattern_code(streaming_pattern, previous_timestep), pdfs, None, None,
                             strain_rate_tensor)
