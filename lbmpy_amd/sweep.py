"""The compiled-sweep callable returned by ``create_lb_function``.

lbmpy's ``create_lb_function`` returns a pystencils kernel that is called as
``kernel(src=array, dst=array, omega=1.8)`` -- fields by their symbolic names,
free parameters (symbolic relaxation rates) by symbol name
(``creationfunctions.py:543-559``, ``lbstep.py:269,273-280``).  ``SweepKernel``
keeps that calling convention for PyTorch ROCm tensors and launches the
MI355X kernel through the C ABI (``include/lbm355.h``).

PDF arrays are indexed like lbmpy's numpy/cupy arrays, ``arr[x, y, (z,) q]``,
with the "fzyx" layout (x contiguous); :func:`create_pdf_array` allocates one.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from .kernel_tables import method_table, family_of, STENCIL_CODES, UnsupportedMethod, bind_rate, \
    cumulant_is_default_set, moment_shear_projector, central_closed_forms, \
    cumulant_corrections
from .relaxationrates import get_shear_relaxation_rate
from .enums import SubgridScaleModel, Target

TORCH_DTYPES = {"float64": torch.float64, "float32": torch.float32, "float16": torch.float16}


def create_pdf_array(domain_size, q, ghost_layers=1, dtype="float64", device="cuda", fill=0.0):
    """Allocate a PDF field in fzyx layout; returns the ``[x, y, (z,) q]`` view."""
    dims = [n + 2 * ghost_layers for n in domain_size]
    storage = torch.full([q] + dims[::-1], fill, dtype=TORCH_DTYPES[str(dtype)], device=device)
    return storage.permute(*range(len(dims), 0, -1), 0)


def create_field_array(domain_size, values_per_cell, ghost_layers=1, dtype="float64", device="cuda", fill=0.0):
    """Macroscopic field (density: values_per_cell=1, velocity: D) in the same layout."""
    arr = create_pdf_array(domain_size, values_per_cell, ghost_layers, dtype, device, fill)
    return arr[..., 0] if values_per_cell == 1 else arr


class UpdateRule:
    """Stand-in for lbmpy's update rule: the method plus the streaming/kernel specification."""

    def __init__(self, method, lbm_config, lbm_optimisation, config):
        self.method = method
        self.lbm_config = lbm_config
        self.lbm_optimisation = lbm_optimisation
        self.config = config

    def __repr__(self):
        c = self.lbm_config
        return (f"UpdateRule({self.method!r}, pattern={c.streaming_pattern}, timestep={c.timestep}, "
                f"dtype={self.config.data_type})")


class KernelAst:
    """Stand-in for the pystencils AST: ``compile()`` yields the GPU sweep."""

    def __init__(self, update_rule):
        self.update_rule = update_rule
        self.method = update_rule.method

    def compile(self):
        return SweepKernel(self.update_rule)


class SweepKernel:
    """One stream-collide step per call, on the current torch stream (asynchronous)."""

    def __init__(self, update_rule):
        self.update_rule = update_rule
        cfg = update_rule.config
        lc = update_rule.lbm_config
        lo = update_rule.lbm_optimisation
        m = update_rule.method
        if cfg.target != Target.GPU:
            raise NotImplementedError("this package implements the GPU (MI355X) sweep; "
                                      "create the kernel with CreateKernelConfig(target=Target.GPU)")
        if lc.kernel_type != 'default_stream_collide':
            raise NotImplementedError(f"kernel_type {lc.kernel_type!r} is not implemented on the GPU path")
        if lc.streaming_pattern not in ("pull", "push") + tuple(nat.INPLACE_EVEN):
            raise ValueError("Invalid value of parameter 'streaming_pattern'.", lc.streaming_pattern)
        if lc.streaming_pattern in nat.INPLACE_EVEN and int(lc.timestep) not in (0, 1):
            raise ValueError(f"Invalid timestep for streaming pattern {lc.streaming_pattern}: {lc.timestep}")
        if lo.field_layout not in ("fzyx", "f", "reverse_numpy", "SoA"):
            raise NotImplementedError("the GPU sweep uses the fzyx (SoA) layout")
        if m.stencil.name not in STENCIL_CODES or not m.stencil.is_gpu_layout():
            raise NotImplementedError(f"stencil {m.stencil.name} ({m.stencil.ordering}) has no GPU sweep")
        self.method = m
        self.lbm_config = lc
        self.config = cfg
        self.family = family_of(m)
        if lc.velocity_input is not None or lc.density_input is not None:
            if lc.velocity_input is None:
                raise ValueError("When setting 'density_input' parameter, 'velocity_input' has to be specified "
                                 "as well.")
            if self.family != nat.FAMILY_MOMENT:
                raise NotImplementedError("density/velocity input fields are implemented for the SRT/TRT/MRT sweep")
        unknown = set(lc.output) - {"density", "velocity"}
        if unknown:
            raise NotImplementedError(f"the GPU sweep writes the outputs 'density' and 'velocity', not {sorted(unknown)}")
        self.store = nat.DTYPE_CODES[cfg.data_type]
        self.compute = nat.DTYPE_CODES[cfg.compute_type]
        if self.store == nat.F16 and self.compute != nat.F32:
            raise ValueError("FP16 storage needs FP32 arithmetic")
        if self.compute not in (nat.F32, nat.F64):
            raise ValueError("compute type must be float32 or float64")
        self.stochastic_rounding = bool(getattr(cfg, "stochastic_rounding", False))
        if self.stochastic_rounding and self.store != nat.F16:
            raise ValueError("stochastic rounding applies to FP16 storage")
        self.rounding_seed = int(getattr(cfg, "rounding_seed", 0)) & 0xFFFFFFFF
        #: launches so far; with stochastic rounding the random stream of a launch depends on (seed, this count,
        #: the kernel's streaming pattern code), so a run restarted with the same count repeats itself bit for bit
        self.rounding_launches = 0
        if self.store == nat.F16 and not self.stochastic_rounding:
            _warn_if_force_is_rounded_away(m)
        self.ghost_layers = int(cfg.ghost_layers)
        if lc.streaming_pattern == "pull":
            self.pattern = nat.PULL
        elif lc.streaming_pattern == "push":
            self.pattern = nat.PUSH
        else:
            self.pattern = nat.INPLACE_EVEN[lc.streaming_pattern] + int(lc.timestep)
        self.two_arrays = self.pattern in (nat.PULL, nat.PUSH)
        self.faces = [nat.FACE_PERIODIC if p else nat.FACE_GHOST for p in lo.builtin_periodicity[:m.dim]]
        self.smagorinsky = None
        sgs = lc.subgrid_scale_model
        if sgs is not None:
            model, const = (sgs if isinstance(sgs, tuple) else (sgs, None))
            if model != SubgridScaleModel.SMAGORINSKY:
                raise NotImplementedError(f"subgrid model {model} is not implemented on the GPU path")
            self.smagorinsky = float(const) if const is not None else 0.12
        if (lc.omega_output_field is not None or lc.eddy_viscosity_field is not None) and self.smagorinsky is None:
            raise NotImplementedError("omega_output_field / eddy_viscosity_field are written by the Smagorinsky "
                                      "model on the GPU path; set subgrid_scale_model")
        #: use the closed-form kernel paths the method admits -- cumulants of order >= 3 at rate 1 set to
        #: equilibrium without being computed, an MRT's shear moments relaxed through their projector in closed
        #: form -- (False runs the generic table-driven forms, for comparisons)
        self.shortcuts = True
        #: a cumulant set other than the default polynomial one: relaxed by the general kernel variant
        self.general_cumulant = self.family == nat.FAMILY_CUMULANT and not cumulant_is_default_set(m)
        if self.general_cumulant and self.smagorinsky is not None:
            raise NotImplementedError("the Smagorinsky model acts on the shear rates of the default polynomial "
                                      "cumulant set")
        self.closure = None
        #: occupancy cap of this kernel's launches: 0 = the library's table, -1 = none, 1..8 = waves per SIMD
        #: (results do not depend on it; tune_occupancy picks the fastest on the GPU at hand)
        self.waves_per_simd = 0
        self._shear_rate = None          # (bindings key, bound shear rate): the lookup is symbolic and slow
        self._tables = {}
        self.corrections = cumulant_corrections(lc)
        if self.corrections is not None and self.corrections.fourth_order:
            if self.smagorinsky is not None:
                raise NotImplementedError("the fourth-order correction is not combined with a subgrid model")
            # the method's rates are the corr_rr symbols; the free ones are those of the shear / bulk rates
            self.free_symbols = self.corrections.free_symbols()
        else:
            self.free_symbols = m.relaxation_rate_symbols()
        # validate the method once with dummy bindings so unsupported configurations fail here
        # (rates of 1.7 / 1.1 keep the fourth-order parametrisation regular)
        probe = {s: (1.7 if k == 0 else 1.1) for k, s in enumerate(self.free_symbols)}
        method_table(m, probe, corrections=self.corrections, inputs=lc.velocity_input is not None)
        nat.load()

    # ------------------------------------------------------------------ configuration
    def set_closure(self, model):
        """Use a learned closure (:class:`lbmpy_amd.closure.ClosureModel`) for the shear rate of this FP32
        cumulant sweep in place of the method's fixed rate (or its Smagorinsky rate); ``None`` removes it."""
        if model is not None:
            if self.family != nat.FAMILY_CUMULANT or self.compute != nat.F32 or self.store != nat.F32:
                raise NotImplementedError("the learned closure runs in the FP32 cumulant sweep with FP32 storage")
            if self.smagorinsky is not None:
                raise ValueError("a sweep uses either the Smagorinsky model or a learned closure")
            if self.corrections is not None:
                raise NotImplementedError("the learned closure is not combined with the cumulant corrections")
            if self.general_cumulant:
                raise NotImplementedError("the learned closure sets the shear rate of the default polynomial "
                                          "cumulant set")
        self.closure = model
        return self

    def set_faces(self, faces):
        """Per-axis face treatment: 'periodic', 'noslip' or 'ghost'."""
        codes = {"periodic": nat.FACE_PERIODIC, "noslip": nat.FACE_NOSLIP, "ghost": nat.FACE_GHOST}
        faces = [codes[f] if isinstance(f, str) else int(f) for f in faces]
        if self.pattern not in (nat.PULL, nat.AA_EVEN, nat.AA_ODD) and nat.FACE_NOSLIP in faces:
            raise ValueError("Eso and push streaming take walls as boundary links, not fused NoSlip faces")
        self.faces = faces
        return self

    #: method tables kept per kernel (one per binding of the free rate symbols and device); the least recently
    #: used one is dropped beyond this, so a loop over time-varying rates holds a bounded amount of GPU memory
    TABLE_CACHE_SIZE = 8

    def table(self, bindings, device):
        key = (tuple(sorted(bindings.items())), str(device))
        t = self._tables.pop(key, None)
        if t is None:
            arr, nact = method_table(self.method, bindings, corrections=self.corrections,
                                     inputs=self.lbm_config.velocity_input is not None)
            dt = torch.float64 if self.compute == nat.F64 else torch.float32
            t = (torch.tensor(arr, dtype=dt, device=device), nact, self._table_flags(arr))
            while len(self._tables) >= self.TABLE_CACHE_SIZE:
                # a launch queued with an older table may still read it: the stream orders the caching
                # allocator's reuse of the freed block after that launch
                self._tables.pop(next(iter(self._tables)))
        self._tables[key] = t
        return t

    def _table_flags(self, arr):
        """The closed-form kernel paths this table admits: LBM355_FLAG_HIGHER_ORDERS_RATE1 for the default D3Q27
        cumulant set whose rates 10..26 (orders 3 to 6) are all 1, without corrections;
        LBM355_FLAG_SHEAR_PROJECTOR for a moment-family table whose individually relaxed moments are the shear
        moments (moment_shear_projector); for a central-moment table, the flags its checked closed forms allow
        (central_closed_forms)."""
        flags = 0
        if (self.family == nat.FAMILY_CUMULANT and not self.general_cumulant and self.corrections is None
                and bool(np.all(np.asarray(arr[10:27]) == 1.0))):
            flags |= nat.FLAG_HIGHER_ORDERS_RATE1
        if self.family == nat.FAMILY_MOMENT and moment_shear_projector(np.asarray(arr), self.method.stencil.Q):
            flags |= nat.FLAG_SHEAR_PROJECTOR
        if self.family == nat.FAMILY_CENTRAL:
            gauss, shear = central_closed_forms(np.asarray(arr))
            if gauss:
                flags |= nat.FLAG_GAUSSIAN_EQUILIBRIUM
            if shear:
                flags |= nat.FLAG_SHEAR_PROJECTOR
        return flags

    def shear_rate(self, bindings):
        """The method's shear relaxation rate with the free symbols bound (cached per binding set)."""
        key = tuple(sorted(bindings.items()))
        if self._shear_rate is None or self._shear_rate[0] != key:
            self._shear_rate = (key, bind_rate(get_shear_relaxation_rate(self.method), bindings))
        return self._shear_rate[1]

    def descriptor(self, arr, bindings, out_strides=(0, 0, 0)):
        m = self.method
        D, Q = m.dim, m.stencil.Q
        g = self.ghost_layers
        if arr.dim() != D + 1 or arr.shape[-1] != Q:
            raise ValueError(f"PDF array must be indexed [x, y{', z' if D == 3 else ''}, q] with q = {Q}")
        if arr.stride(0) != 1:
            raise ValueError("PDF array must be in fzyx layout (x contiguous)")
        sizes = [arr.shape[a] - 2 * g for a in range(D)]
        if min(sizes) <= 0:
            raise ValueError("PDF array smaller than its ghost layers")
        tab, nact, table_flags = self.table(bindings, arr.device)
        d = nat.SweepDesc()
        d.stencil = STENCIL_CODES[m.stencil.name]
        d.family = self.family
        d.pattern = self.pattern
        d.store_dtype = self.store
        d.compute_dtype = self.compute
        d.nx, d.ny = sizes[0], sizes[1]
        d.nz = sizes[2] if D == 3 else 1
        d.ghost_layers = g
        for a in range(3):
            d.face[a] = self.faces[a] if a < D else nat.FACE_PERIODIC
        d.stride_y = arr.stride(1)
        d.stride_z = arr.stride(2) if D == 3 else 0
        d.stride_q = arr.stride(D)
        d.out_stride_y, d.out_stride_z, d.out_stride_c = out_strides
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
        if self.smagorinsky is not None:
            flags |= nat.FLAG_SMAGORINSKY
            d.smagorinsky_constant = self.smagorinsky
            d.omega_shear = self.shear_rate(bindings)
        if self.corrections is not None:
            if self.corrections.galilean:
                flags |= nat.FLAG_GALILEAN
            if self.corrections.fourth_order:
                flags |= nat.FLAG_FOURTH_ORDER
        if self.general_cumulant:
            flags |= nat.FLAG_GENERAL_CUMULANT
        if self.shortcuts:
            flags |= table_flags
        if self.closure is not None:
            flags |= nat.FLAG_LEARNED_CLOSURE
            d.closure = self.closure.params.data_ptr()
            d.omega_shear = self.shear_rate(bindings)
        d.flags = flags
        d.n_active = nact
        d.tables = tab.data_ptr()
        d.waves_per_simd = int(self.waves_per_simd)
        return d

    # ------------------------------------------------------------------ call
    def __call__(self, **kwargs):
        lc = self.lbm_config
        src = kwargs.pop(lc.field_name)
        dst = kwargs.pop(lc.temporary_field_name, None)
        if self.two_arrays:
            if dst is None:
                raise ValueError(f"{lc.streaming_pattern} streaming needs the temporary field "
                                 f"'{lc.temporary_field_name}'")
        else:
            dst = src
        bindings = {}
        for s in self.free_symbols:
            if s not in kwargs:
                raise ValueError(f"missing value for free symbol '{s}'")
            bindings[s] = float(kwargs.pop(s))

        def field(target):
            # a tensor given in the config, or the name under which the call passes it (a pystencils field's name)
            if target is None:
                return None
            if isinstance(target, str):
                return kwargs.pop(target, None)
            name = getattr(target, "name", None)
            if not isinstance(target, torch.Tensor) and isinstance(name, str):
                return kwargs.pop(name, None)
            return target
        fields = {"density": field(lc.output.get("density")), "velocity": field(lc.output.get("velocity")),
                  "omega": field(lc.omega_output_field), "eddy_viscosity": field(lc.eddy_viscosity_field),
                  "density_input": field(lc.density_input), "velocity_input": field(lc.velocity_input)}
        self._check_tensor(src)
        self._check_tensor(dst)
        if self.two_arrays and src.shape != dst.shape:
            raise ValueError("src and dst must have the same shape")
        if self.two_arrays and src.stride() != dst.stride():
            raise ValueError("src and dst must have the same strides")
        if (lc.velocity_input is not None) != (fields["velocity_input"] is not None) or \
                (lc.density_input is not None) != (fields["density_input"] is not None):
            raise ValueError("the density / velocity input fields are missing from the call")
        out_strides = self._cell_field_strides(src, fields)
        d = self.descriptor(src, bindings, out_strides)
        if self.stochastic_rounding:
            d.flags |= nat.FLAG_STOCHASTIC_ROUNDING
            d.rounding_seed = launch_seed(self.rounding_seed, self.rounding_launches, self.pattern)
            self.rounding_launches += 1
        ptr = (lambda t: t.data_ptr() if t is not None else None)
        d.omega_out = ptr(fields["omega"])
        d.eddy_viscosity_out = ptr(fields["eddy_viscosity"])
        d.density_in = ptr(fields["density_input"])
        d.velocity_in = ptr(fields["velocity_input"])
        lib = nat.load()
        nat.check(lib.lbm355_sweep(ctypes.byref(d), ctypes.c_void_p(src.data_ptr()),
                                   ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(ptr(fields["density"])),
                                   ctypes.c_void_p(ptr(fields["velocity"])), nat.current_stream_handle()))

    def _cell_field_strides(self, src, fields):
        """(stride_y, stride_z, component stride) shared by the per-cell fields beside the PDFs, after checking
        that each is a compute-type tensor on the PDFs' device, indexed [x, y, (z,) (component)] over the PDF
        array's spatial extent with x contiguous: the kernel writes and reads every cell of the sweep there."""
        D = self.method.dim
        ct = torch.float64 if self.compute == nat.F64 else torch.float32
        given = {k: t for k, t in fields.items() if t is not None}
        if not given:
            return (0, 0, 0)
        spatial = tuple(src.shape[:D])
        strides = None
        comp = None
        for name, t in given.items():
            vector = name in ("velocity", "velocity_input")
            want = spatial + ((D,) if vector else ())
            if not isinstance(t, torch.Tensor) or t.device != src.device or t.dtype != ct:
                raise ValueError(f"the {name} field must be a {ct} tensor on {src.device}")
            if tuple(t.shape) != want:
                raise ValueError(f"the {name} field must have shape {want} (the PDF array's spatial shape"
                                 f"{' and D components' if vector else ''}), not {tuple(t.shape)}")
            if t.stride(0) != 1:
                raise ValueError(f"the {name} field must be contiguous in x (fzyx layout)")
            if strides is None:
                strides = t.stride()[:D]
            elif t.stride()[:D] != strides:
                raise ValueError("the per-cell fields (outputs and inputs) must share their spatial strides")
            if vector:
                if comp is None:
                    comp = t.stride(D)
                elif t.stride(D) != comp:
                    raise ValueError("the velocity output and input must share their component stride")
        volume = max((spatial[a] - 1) * strides[a] for a in range(D)) + 1
        if comp is not None and comp < volume:
            raise ValueError("velocity components must not overlap (fzyx layout: component slowest)")
        return (strides[1], strides[2] if D == 3 else 0, comp if comp is not None else 0)

    def conserved(self, pdf, **bindings):
        """FP64 sums over the interior on the GPU: (mass, momentum_x, momentum_y, momentum_z) of the stored
        populations, read in their storage slots (for AA fields only the mass is layout independent).
        Zero-centred storage counts the rest density 1 per cell.  Synchronises the current stream."""
        self._check_tensor(pdf)
        binds = {s: float(bindings.get(s, 1.0)) for s in self.free_symbols}
        d = self.descriptor(pdf, binds)
        out = torch.zeros(4, dtype=torch.float64, device=pdf.device)
        lib = nat.load()
        nat.check(lib.lbm355_reduce_conserved(ctypes.byref(d), ctypes.c_void_p(pdf.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr()), nat.current_stream_handle()))
        return out.cpu().numpy()

    def nonfinite_count(self, pdf, **bindings):
        """Number of interior populations of `pdf` that are NaN or infinite, counted on the GPU
        (lbm355_count_nonfinite); 0 for a healthy field.  Synchronises the current stream."""
        self._check_tensor(pdf)
        binds = {s: float(bindings.get(s, 1.0)) for s in self.free_symbols}
        d = self.descriptor(pdf, binds)
        out = torch.zeros(1, dtype=torch.int64, device=pdf.device)
        lib = nat.load()
        nat.check(lib.lbm355_count_nonfinite(ctypes.byref(d), ctypes.c_void_p(pdf.data_ptr()),
                                             ctypes.c_void_p(out.data_ptr()), nat.current_stream_handle()))
        return int(out.item())

    def _check_tensor(self, t):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError("PDF fields must be torch tensors on the GPU")
        if t.dtype != TORCH_DTYPES[self.config.data_type]:
            raise TypeError(f"PDF field dtype {t.dtype} != kernel data type {self.config.data_type}")

    @property
    def parameters(self):
        return list(self.free_symbols)

    def __repr__(self):
        return f"SweepKernel({self.update_rule!r})"


def tune_occupancy(kernels, src, dst=None, candidates=(0, -1, 1, 2, 3, 4, 5, 6), reps=8, **params):
    """Time every kernel of a time loop at each occupancy cap in `candidates` (``SweepKernel.waves_per_simd``: 0 the
    library's table, -1 none, n at most n waves per SIMD) and leave each kernel at its fastest; returns
    {kernel index: {cap: ms per launch}}.

    `kernels` are the loop's kernels in launch order: one pull / push kernel (which swaps `src` and `dst`), or the
    ODD and EVEN kernels of an in-place pattern (which alternate on `src`).  The loop runs on copies of the arrays
    (GPU memory for one more set of them), so the caller's state does not advance; `params` are the kernels'
    free-symbol values.  The best cap of a kernel
    can differ from one MI355X to another: the caps of the library's table were measured on some and are one or two
    waves off on others."""
    import torch
    a = src.clone()
    b = dst.clone() if dst is not None else None
    # stochastic rounding draws a new seed per launch: the tuning launches must not shift the caller's sequence
    launches = [getattr(k, "rounding_launches", None) for k in kernels]

    def launch(j, k):
        names = k.lbm_config.field_name, k.lbm_config.temporary_field_name
        if b is not None:
            k(**{names[0]: a if j % 2 == 0 else b, names[1]: b if j % 2 == 0 else a}, **params)
        else:
            k(**{names[0]: a}, **params)

    def rounds(count, timed=None, ev=None):
        for r in range(count):
            for j, kk in enumerate(kernels):
                if j == timed:
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    launch(len(kernels) * r + j, kk)
                    e1.record()
                    ev.append((e0, e1))
                else:
                    launch(len(kernels) * r + j, kk)

    rounds(4)                                   # the GPU's clocks settle before the first candidate is timed
    result = {}
    for i, k in enumerate(kernels):
        times = {}
        keep = k.waves_per_simd
        # two passes over the candidates; each keeps its better pass, which damps the drift between them
        for _ in range(2):
            for cap in candidates:
                k.waves_per_simd = cap
                rounds(1)
                ev = []
                rounds(reps, i, ev)
                torch.cuda.synchronize()
                t = sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev)
                times[cap] = min(t, times.get(cap, t))
        k.waves_per_simd = min(times, key=times.get) if times else keep
        result[i] = times
    for k, n in zip(kernels, launches):
        if n is not None:
            k.rounding_launches = n
    del a, b
    return result


def _mix32(x):
    """A 32-bit integer hash (xorshift-multiply), the same mixing the kernels use for their rounding bits."""
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def launch_seed(seed, launch, pattern):
    """Seed of one stochastically rounded launch: distinct per (seed, launch count, streaming-pattern code), so
    the two kernels of an in-place pattern never share a random stream."""
    return _mix32((seed * 0x9E3779B1 + launch * 0x85EBCA77 + pattern * 0xC2B2AE3D + 0x165667B1) & 0xFFFFFFFF)


#: lattice velocity at which the FP16 ulp of the populations is judged against a force increment
FP16_REFERENCE_VELOCITY = 0.01


def force_rounded_away_in_fp16(method):
    """True when round-to-nearest FP16 storage drops the body force: for every direction with c.F != 0 the per-step
    force increment 3 w |c.F| is below half an FP16 ulp of a zero-centred population at lattice velocity 0.01
    (3 w 0.01).  Such a force only survives with stochastic rounding (or FP32 storage)."""
    fm = method.force_model
    if fm is None:
        return False
    F = np.zeros(3)
    F[:method.dim] = [float(v) for v in fm.force]
    c = np.array([tuple(e) + (0,) * (3 - method.dim) for e in method.stencil], dtype=np.float64)
    w = np.array([float(v) for v in method.weights])
    incr = 3.0 * w * np.abs(c @ F)
    sel = incr > 0
    if not sel.any():
        return False
    dev = (3.0 * w[sel] * FP16_REFERENCE_VELOCITY).astype(np.float16)
    half_ulp = 0.5 * np.spacing(dev).astype(np.float64)
    return bool(np.all(incr[sel] < half_ulp))


def _warn_if_force_is_rounded_away(method):
    if force_rounded_away_in_fp16(method):
        import warnings
        warnings.warn("FP16 storage with round-to-nearest drops this body force: its per-step increment on every "
                      "population is below half an FP16 ulp; use CreateKernelConfig(stochastic_rounding=True) or "
                      "FP32 storage", RuntimeWarning, stacklevel=3)
