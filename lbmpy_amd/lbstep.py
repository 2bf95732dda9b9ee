"""``LatticeBoltzmannStep`` on the MI355X sweep.

Mirrors lbmpy ``src/lbmpy/lbstep.py:18-450``: same constructor parameters and
defaults (D2Q9 / D3Q27 by dimension, one ghost layer, ``<name>_pdfSrc`` /
``<name>_pdfTmp`` fields), ``run``, ``time_step``, ``pre_run``/``post_run``,
``benchmark_run``/``benchmark`` (MLUPS = cells / seconds-per-step * 1e-6,
lbstep.py:329-344), ``velocity``/``density`` slicing, ``boundary_handling``
with ``set_boundary(NoSlip(), slice_from_direction(...))`` and
``set_pdf_fields_from_macroscopic_values``.

Differences, all deliberate: the target is always the GPU (this package has
no CPU sweep); PDF fields are PyTorch tensors that stay resident in HBM; the
macroscopic setter/getter run on the GPU (the reference compiles them for the
CPU, lbstep.py:444-449); periodicity is built into the sweep instead of ghost
copies; besides pull, the time loop runs push and the in-place patterns 'aa',
'esotwist', 'esopull' and 'esopush' (EVEN/ODD kernel pairs, ODD first after
the setter, boundaries between the kernels as lbmpy's BetweenTimestepsIndexing
addresses them), which the reference's fixed time loop does not support.
"""
import ctypes
import time
from types import MappingProxyType

import numpy as np
import torch

from . import _native as nat
from .boundaries import LatticeBoltzmannBoundaryHandling
from .creationfunctions import (update_with_default_parameters, create_lb_function, Timestep, CreateKernelConfig)
from .enums import Stencil, Target
from .kernel_tables import method_table
from .stencils import LBStencil
from .sweep import create_pdf_array, TORCH_DTYPES


class _SlicedGetter:
    def __init__(self, fn):
        self._fn = fn

    def __getitem__(self, item):
        return self._fn(item)


class LatticeBoltzmannStep:

    def __init__(self, domain_size=None, lbm_kernel=None, periodicity=False, kernel_params=MappingProxyType({}),
                 data_handling=None, name="lbm", optimization=None, velocity_data_name=None, density_data_name=None,
                 density_data_index=None, compute_velocity_in_every_step=False, compute_density_in_every_step=False,
                 velocity_input_array_name=None, time_step_order='stream_collide', flag_interface=None,
                 alignment_if_vectorized=64, fixed_loop_sizes=True, timeloop_creation_function=None,
                 lbm_config=None, lbm_optimisation=None, config=None, **method_parameters):
        if data_handling is not None:
            raise NotImplementedError("pystencils data handlings are not used on the GPU path; pass domain_size")
        if domain_size is None:
            raise ValueError("Specify either domain_size or data_handling")
        if time_step_order != 'stream_collide':
            raise NotImplementedError("only the fused stream-collide order is implemented")
        if velocity_input_array_name is not None:
            raise NotImplementedError("velocity input arrays are not supported")
        optimization = dict(optimization or {})
        optimization["target"] = Target.GPU
        domain_size = tuple(int(n) for n in domain_size)
        dim = len(domain_size)
        if lbm_config:
            method_parameters['stencil'] = lbm_config.stencil
        if 'stencil' not in method_parameters:
            method_parameters['stencil'] = LBStencil(Stencil.D2Q9 if dim == 2 else Stencil.D3Q27)
        if config is not None:
            from dataclasses import replace
            config = replace(config, target=Target.GPU)
        lbm_config, lbm_optimisation, config = update_with_default_parameters(
            method_parameters, optimization, lbm_config, lbm_optimisation, config)
        if isinstance(periodicity, bool):
            periodicity = (periodicity,) * dim
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.name = name
        self._pdf_arr_name = name + "_pdfSrc"
        self._tmp_arr_name = name + "_pdfTmp"
        self.velocity_data_name = name + "_velocity" if velocity_data_name is None else velocity_data_name
        self.density_data_name = name + "_density" if density_data_name is None else density_data_name
        self.kernel_params = dict(kernel_params)
        self._domain_size = domain_size
        self._lbm_config = lbm_config
        self._lbm_optimisation = lbm_optimisation
        self._config = config
        self._pattern = lbm_config.streaming_pattern
        if self._pattern not in ("pull", "push") + tuple(nat.INPLACE_EVEN):
            raise ValueError("Invalid value of parameter 'streaming_pattern'.", self._pattern)
        self._two_arrays = self._pattern in ("pull", "push")
        # kernel codes: (the single kernel,) for pull / push, (EVEN, ODD) for in-place patterns
        self._codes = ((nat.PULL,) if self._pattern == "pull" else (nat.PUSH,)) if self._two_arrays else \
            (nat.INPLACE_EVEN[self._pattern], nat.INPLACE_EVEN[self._pattern] + 1)

        from dataclasses import replace
        lc = replace(lbm_config, field_name=self._pdf_arr_name, temporary_field_name=self._tmp_arr_name,
                     lb_method=None, collision_rule=None, update_rule=None, ast=None)
        cfg = replace(config, ghost_layers=1)
        if lbm_kernel is not None:
            self._kernels = [lbm_kernel]
        elif self._two_arrays:
            self._kernels = [create_lb_function(lbm_config=lc, lbm_optimisation=lbm_optimisation, config=cfg)]
        else:
            self._kernels = [create_lb_function(lbm_config=replace(lc, timestep=t), lbm_optimisation=lbm_optimisation,
                                                config=cfg) for t in (Timestep.EVEN, Timestep.ODD)]
        self.method = self._kernels[0].method
        self.ast = self._kernels[0].ast
        q = self.method.stencil.Q
        dt = config.data_type
        self._src = create_pdf_array(domain_size, q, 1, dt)
        self._tmp = create_pdf_array(domain_size, q, 1, dt) if self._two_arrays else None
        # macroscopic fields: host numpy (interior) + FP64 device staging with the PDF ghost layout
        self._rho_host = np.ones(domain_size)
        self._u_host = np.zeros(domain_size + (dim,))
        self._rho_dev = torch.zeros([n + 2 for n in domain_size][::-1], dtype=torch.float64, device="cuda")
        self._u_dev = torch.zeros([dim] + [n + 2 for n in domain_size][::-1], dtype=torch.float64, device="cuda")
        self._boundary_handling = LatticeBoltzmannBoundaryHandling.for_domain(self.method, domain_size,
                                                                             self.periodicity)
        # full NoSlip planes are fused into the pull / AA sweeps; the other patterns take them as links
        self._boundary_handling.fuse_noslip_faces = self._pattern in ("pull", "aa")
        self._aa_next_odd = True
        self.time_steps_run = 0
        self._graph = None
        self.set_pdf_fields_from_macroscopic_values()

    # ------------------------------------------------------------------ properties
    @property
    def boundary_handling(self):
        return self._boundary_handling

    @property
    def data_handling(self):
        return None

    @property
    def dim(self):
        return len(self._domain_size)

    @property
    def domain_size(self):
        return self._domain_size

    @property
    def number_of_cells(self):
        return int(np.prod(self._domain_size))

    @property
    def pdf_array_name(self):
        return self._pdf_arr_name

    @property
    def lbm_config(self):
        return self._lbm_config

    @property
    def lbm_optimisation(self):
        return self._lbm_optimisation

    @property
    def config(self):
        return self._config

    @property
    def pdf_array(self):
        """The current PDF field (torch, ``[x, y, (z,) q]`` with one ghost layer)."""
        return self._src

    @property
    def velocity(self):
        return _SlicedGetter(lambda s: self.velocity_slice(s, masked=False))

    @property
    def density(self):
        return _SlicedGetter(lambda s: self.density_slice(s, masked=False))

    # ------------------------------------------------------------------ fields
    def _default_slice(self):
        return (slice(None),) * 2 if self.dim == 2 else (slice(None), slice(None), self._domain_size[2] // 2)

    def velocity_slice(self, slice_obj=None, masked=True):
        res = self._u_host[self._default_slice() if slice_obj is None else slice_obj]
        if masked:
            mask = self._boundary_handling.get_mask(None, 'domain', True)
            sl = self._default_slice() if slice_obj is None else slice_obj
            m = mask[sl]
            if m.ndim < res.ndim:
                m = np.repeat(m[..., None], res.shape[-1], axis=-1)
            res = np.ma.masked_array(res, m)
        return res

    def density_slice(self, slice_obj=None, masked=True):
        res = self._rho_host[self._default_slice() if slice_obj is None else slice_obj]
        if masked:
            mask = self._boundary_handling.get_mask(None, 'domain', True)
            res = np.ma.masked_array(res, mask[self._default_slice() if slice_obj is None else slice_obj])
        return res

    def _descriptor(self, arr=None):
        k = self._kernels[0]
        k.set_faces(self._boundary_handling.faces)
        d = k.descriptor(self._src if arr is None else arr, self.kernel_params)
        D = self.dim
        d.out_stride_y = self._rho_dev.stride(D - 2)
        d.out_stride_z = self._rho_dev.stride(0) if D == 3 else 0
        d.out_stride_c = self._u_dev.stride(0)
        return d

    def set_pdf_fields_from_macroscopic_values(self):
        """Equilibrium from the host density/velocity fields (lbstep.py:268-269), on the GPU."""
        D = self.dim
        inner = (slice(1, -1),) * D
        self._rho_dev[inner] = torch.from_numpy(np.ascontiguousarray(self._rho_host.T)).cuda()
        for a in range(D):
            self._u_dev[a][inner] = torch.from_numpy(np.ascontiguousarray(self._u_host[..., a].T)).cuda()
        d = self._descriptor()
        tab, _ = method_table(self.method, self.kernel_params, np.float64,
                              corrections=self._kernels[0].corrections)
        tab_dev = torch.from_numpy(tab).cuda()
        prev = self._codes[0]
        lib = nat.load()
        nat.check(lib.lbm355_set_equilibrium(ctypes.byref(d), ctypes.c_void_p(self._src.data_ptr()),
                                             ctypes.c_void_p(self._rho_dev.data_ptr()),
                                             ctypes.c_void_p(self._u_dev.data_ptr()),
                                             ctypes.c_void_p(tab_dev.data_ptr()), prev, nat.current_stream_handle()))
        self._aa_next_odd = True
        torch.cuda.synchronize()

    def _apply_links(self, pattern):
        self._boundary_handling.apply(self._src, pattern)

    # ------------------------------------------------------------------ stepping
    def pre_run(self):
        for k in self._kernels:
            k.set_faces(self._boundary_handling.faces)

    def time_step(self):
        if self._two_arrays:
            self._apply_links(self._codes[0])
            self._kernels[0](**{self._pdf_arr_name: self._src, self._tmp_arr_name: self._tmp}, **self.kernel_params)
            self._src, self._tmp = self._tmp, self._src
        else:
            even, odd = self._kernels
            if self._aa_next_odd:
                self._apply_links(self._codes[0])
                odd(**{self._pdf_arr_name: self._src}, **self.kernel_params)
            else:
                self._apply_links(self._codes[1])
                even(**{self._pdf_arr_name: self._src}, **self.kernel_params)
            self._aa_next_odd = not self._aa_next_odd

    def post_run(self):
        """Density and velocity of the latest populations into the host fields (lbstep.py:311-314)."""
        d = self._descriptor()
        if self._two_arrays:
            last = self._codes[0]
        else:
            last = self._codes[0] if self._aa_next_odd else self._codes[1]
        lib = nat.load()
        nat.check(lib.lbm355_get_macroscopic(ctypes.byref(d), ctypes.c_void_p(self._src.data_ptr()),
                                             ctypes.c_void_p(self._rho_dev.data_ptr()),
                                             ctypes.c_void_p(self._u_dev.data_ptr()), last,
                                             nat.current_stream_handle()))
        D = self.dim
        inner = (slice(1, -1),) * D
        self._rho_host[...] = self._rho_dev[inner].cpu().numpy().T
        for a in range(D):
            self._u_host[..., a] = self._u_dev[a][inner].cpu().numpy().T

    # ------------------------------------------------------------------ fixed two-step loop
    # lbmpy's run() and benchmark_run() go through a time loop of two fixed steps with the kernel arguments bound
    # once (lbstep.py:283-309, 316-327).  Here the two steps (boundary links and sweep, twice) are captured into a
    # HIP graph once and replayed, so that a step launches no Python; the state the captured launches depend on is
    # the graph's key, and any change to it captures again.  Runs whose launches change from step to step
    # (stochastic rounding seeds) and outflow boundaries before their first application run step by step.
    # On MI355X (D3Q19 TRT cavity, FP32): 32^3 830 -> 1,825 MLUPS, 64^3 7,579 -> 9,307 MLUPS; from about 96^3 on
    # the launches queue faster than the GPU runs them and a graph gains nothing (128^3: 21,273 vs 20,842).
    #: replay captured two-step graphs in run() and benchmark_run(); False runs every step from Python
    use_graphs = True
    #: domains of at least this many cells run step by step (their launches already keep the GPU busy)
    graph_max_cells = 1 << 20

    def _graph_key(self):
        """What the two captured steps depend on, or None when they cannot be replayed from a graph."""
        if not self.use_graphs or self.number_of_cells >= self.graph_max_cells or \
                any(getattr(k, "stochastic_rounding", False) for k in self._kernels):
            return None
        bh = self._boundary_handling
        faces = tuple(bh.faces)
        groups = bh._device_groups(self._src.device)
        if not all(g[3] for g in groups):
            return None
        bparams = tuple((type(o).__name__, id(l), id(v), repr(getattr(o, "_velocity", None)),
                         repr(getattr(o, "_density", None)), repr(getattr(o, "density", None)),
                         repr(getattr(o, "normal_direction", None))) for o, l, v, _ in groups)
        params = tuple(sorted((str(k), repr(v)) for k, v in self.kernel_params.items()))
        kernels = tuple((id(k), id(getattr(k, "closure", None)), tuple(getattr(k, "faces", None) or ()),
                         getattr(k, "waves_per_simd", 0)) for k in self._kernels)
        return (self._src.data_ptr(), self._tmp.data_ptr() if self._tmp is not None else 0, self._aa_next_odd,
                faces, bh.version, bparams, params, kernels)

    def _two_step_graph(self, key):
        """The captured two-step graph for `key`, capturing it now if needed (after two steps run from Python, so
        that every lazily built table and link list exists before the capture); returns (graph, eager steps)."""
        if self._graph is not None and self._graph[0] == key:
            return self._graph[1], 0
        self._graph = None
        self.time_step()
        self.time_step()
        key = self._graph_key()
        if key is None:
            return None, 2
        # keep every device buffer the captured launches read alive as long as the graph
        keep = [list(getattr(k, "_tables", {}).values()) for k in self._kernels]
        keep.append(list(self._boundary_handling._device_groups(self._src.device)))
        keep.append(getattr(self._boundary_handling, "_fd_table", None))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.time_step()
            self.time_step()
        # two steps return the array pair and the AA parity to where they were: the Python state is unchanged
        self._graph = (key, g, keep)
        return g, 2

    def _advance(self, time_steps):
        left = time_steps
        while left >= 2:
            key = self._graph_key()
            if key is None:
                self.time_step()
                left -= 1
                continue
            g, eager = self._two_step_graph(key)
            left -= eager
            if g is None:
                continue
            n = left // 2
            for _ in range(n):
                g.replay()
            left -= 2 * n
        for _ in range(left):
            self.time_step()

    def tune_occupancy(self, candidates=(0, -1, 1, 2, 3, 4, 5, 6), reps=8):
        """Pick each sweep kernel's occupancy cap by timing it on a copy of the populations
        (lbmpy_amd.sweep.tune_occupancy); the run's state does not advance.  Returns the timings."""
        from .sweep import tune_occupancy
        self.pre_run()
        self._graph = None
        even_odd = self._kernels
        kernels = list(even_odd) if self._two_arrays else \
            ([even_odd[1], even_odd[0]] if self._aa_next_odd else list(even_odd))
        return tune_occupancy(kernels, self._src, self._tmp if self._two_arrays else None, candidates=candidates,
                              reps=reps, **self.kernel_params)

    def run(self, time_steps):
        self.pre_run()
        self._advance(time_steps)
        self.post_run()
        self.time_steps_run += time_steps

    # ------------------------------------------------------------------ failure detection, checkpoints
    def nonfinite_count(self):
        """Number of interior populations that are NaN or infinite (counted on the GPU); 0 for a healthy run."""
        return self._kernels[0].nonfinite_count(self._src, **self.kernel_params)

    def check_finite(self):
        """Raise FloatingPointError when the populations hold NaN or infinite values."""
        n = self.nonfinite_count()
        if n:
            raise FloatingPointError(f"{n} populations are NaN or infinite after {self.time_steps_run} time steps")

    def save_checkpoint(self, path):
        """Write the run's state: populations, steps run, which AA kernel runs next and the kernels' launch counts
        (lbmpy_amd.checkpoint)."""
        from .checkpoint import save_pdf_state
        torch.cuda.synchronize()
        save_pdf_state(path, self._src, time_step=self.time_steps_run, stencil_name=self.method.stencil.name,
                       pattern=self._pattern, kernels=self._kernels, next_odd=self._aa_next_odd)

    def load_checkpoint(self, path):
        """Resume from a state written by save_checkpoint: same domain, stencil, storage type and streaming
        pattern; the boundary handling and kernel parameters are this object's own."""
        from .checkpoint import load_pdf_state
        self.time_steps_run, self._aa_next_odd = load_pdf_state(path, self._src, stencil_name=self.method.stencil.name,
                                                                pattern=self._pattern, kernels=self._kernels)
        torch.cuda.synchronize()
        self.post_run()

    def benchmark_run(self, time_steps, number_of_cells=None):
        if number_of_cells is None:
            number_of_cells = self.number_of_cells
        self.pre_run()
        if self.use_graphs and time_steps >= 2:
            key = self._graph_key()
            if key is not None:
                g, eager = self._two_step_graph(key)   # capture (and its two steps) before the timed region
                self.time_steps_run += eager
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self._advance(time_steps)
        torch.cuda.synchronize()
        duration = (time.perf_counter() - t0) / time_steps
        self.time_steps_run += time_steps
        self.post_run()
        return number_of_cells / duration * 1e-6

    def benchmark(self, time_for_benchmark=5, init_time_steps=2, number_of_time_steps_for_estimation='auto'):
        self.pre_run()
        for _ in range(init_time_steps):
            self.time_step()
        n_est = 20 if number_of_time_steps_for_estimation == 'auto' else number_of_time_steps_for_estimation
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self._advance(n_est)
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / n_est
        steps = max(2, int(time_for_benchmark / max(per, 1e-9)))
        steps += steps % 2
        return self.benchmark_run(steps)

    def write_vtk(self):
        raise NotImplementedError("VTK output is not part of the GPU path")

    def run_iterative_initialization(self, *args, **kwargs):
        raise NotImplementedError("iterative initialisation is not implemented on the GPU path")
