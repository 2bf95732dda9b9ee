"""Batched ensembles: many independent LB domains advanced by one kernel launch per time step.

lbmpy runs an ensemble as a Python loop over members, each calling its own compiled sweep
(``LatticeBoltzmannStep.time_step``, ``lbstep.py:271-282``).  :class:`BatchedSweep` takes the same
per-member sweep kernels (:func:`lbmpy_amd.create_lb_function` results, one per member because
relaxation rates and forces differ) and the members' PDF arrays, and advances all of them with one
persistent launch (``lbm355_batch_sweep``).  Each member's cells run the same per-cell code as
calling its kernel alone, so results are bit-identical to single-domain runs.

:meth:`BatchedSweep.autotune` picks the launch shape (threads per block, persistent grid size,
XCD-ordered tiles) by timing the candidates on the ensemble itself, and
:meth:`BatchedSweep.capture` records many time steps into one HIP graph, replayed as one launch.
"""
import ctypes
import json
import os

import torch

from . import _native as nat

DEFAULT_CANDIDATES = [(t, m, x) for t in (64, 128, 256) for m in (2, 4, 8, 16, 0) for x in (0, 1)]


def _num_cus(device):
    return torch.cuda.get_device_properties(device).multi_processor_count


class BatchedSweep:
    """One launch per time step over members ``kernels[k]`` acting on ``src[k]`` (and ``dst[k]`` for pull).

    ``bindings``: optional per-member dicts of free relaxation-rate symbols.  The member arrays must
    stay alive (and must not be reallocated) while the batch exists; pull members are swapped
    in place: step 0 reads ``src`` and writes ``dst``, step 1 the reverse, and so on.
    """

    def __init__(self, kernels, src, dst=None, bindings=None):
        if not kernels:
            raise ValueError("empty ensemble")
        n = len(kernels)
        if len(src) != n or (dst is not None and len(dst) != n):
            raise ValueError("one PDF array (two for pull) per member")
        self.kernels = list(kernels)
        self.src = list(src)
        self.pull = kernels[0].pattern == nat.PULL
        if self.pull and dst is None:
            raise ValueError("pull streaming needs a destination array per member")
        self.dst = list(dst) if self.pull else list(src)
        bindings = bindings or [{}] * n
        descs = (nat.SweepDesc * n)()
        srcp = (ctypes.c_void_p * n)()
        dstp = (ctypes.c_void_p * n)()
        for k, (kern, s, d, b) in enumerate(zip(self.kernels, self.src, self.dst, bindings)):
            kern._check_tensor(s)
            kern._check_tensor(d)
            binds = {sym: float(b[sym]) for sym in kern.free_symbols}
            descs[k] = kern.descriptor(s, binds)
            srcp[k] = s.data_ptr()
            dstp[k] = d.data_ptr()
        self.lib = nat.load()
        handle = ctypes.c_void_p()
        nat.check(self.lib.lbm355_batch_create(descs, srcp, dstp, n, ctypes.byref(handle)))
        self._handle = handle
        nt, nc = ctypes.c_int64(), ctypes.c_int64()
        nat.check(self.lib.lbm355_batch_info(handle, ctypes.byref(nt), ctypes.byref(nc)))
        self.n_tiles, self.n_cells = nt.value, nc.value
        self.device = self.src[0].device
        self.launch = (256, 0, 0)          # (threads, blocks per CU (0: one block per tile), xcd order)
        self.steps_done = 0

    # ------------------------------------------------------------------ launching
    def _blocks(self, threads, per_cu):
        if per_cu == 0:
            return (self.n_tiles * 64 + threads - 1) // threads
        return _num_cus(self.device) * per_cu

    def sweep(self, step=None, launch=None):
        """Advance every member by one time step (asynchronous on the current stream)."""
        t = self.steps_done if step is None else step
        threads, per_cu, xcd = launch or self.launch
        if self.pull:
            kind = t % 2
        else:
            kind = nat.AA_ODD if t % 2 == 0 else nat.AA_EVEN       # the first AA step is ODD
        nat.check(self.lib.lbm355_batch_sweep(self._handle, kind, threads, self._blocks(threads, per_cu), xcd,
                                              nat.current_stream_handle()))
        if step is None:
            self.steps_done += 1

    def run(self, steps):
        for _ in range(steps):
            self.sweep()

    def current_arrays(self):
        """The arrays holding each member's latest populations (pull alternates src/dst)."""
        if self.pull and self.steps_done % 2 == 1:
            return self.dst
        return self.src

    # ------------------------------------------------------------------ autotuning
    def autotune(self, candidates=None, reps=6, cache=None):
        """Time each launch shape over `reps` time steps on the ensemble (advancing it by a multiple
        of 2 steps per candidate) and keep the fastest.  ``cache``: optional JSON file keyed by the
        ensemble's signature.  Returns {"launch": (threads, blocks_per_cu, xcd), "ms": {...}}."""
        key = self.signature()
        if cache and os.path.exists(cache):
            with open(cache) as fh:
                table = json.load(fh)
            if key in table:
                self.launch = tuple(table[key]["launch"])
                return table[key]
        reps += reps % 2
        timings = {}
        t0 = self.steps_done
        for cand in candidates or DEFAULT_CANDIDATES:
            for w in range(2):
                self.sweep(step=t0 + w, launch=cand)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for r in range(reps):
                self.sweep(step=t0 + r, launch=cand)
            e1.record()
            torch.cuda.synchronize()
            timings[cand] = e0.elapsed_time(e1) / reps
        best = min(timings, key=timings.get)
        self.launch = best
        result = {"launch": list(best), "ms": {f"{t}x{m}x{x}": round(v, 5) for (t, m, x), v in timings.items()}}
        if cache:
            table = {}
            if os.path.exists(cache):
                with open(cache) as fh:
                    table = json.load(fh)
            table[key] = result
            with open(cache, "w") as fh:
                json.dump(table, fh, indent=1)
        return result

    def signature(self):
        k0 = self.kernels[0]
        sizes = sorted(tuple(s.shape[:-1]) for s in self.src)
        return json.dumps([k0.method.stencil.name, k0.family, k0.pattern, k0.store, k0.compute, len(sizes),
                           self.n_cells, hash(tuple(sizes)) & 0xFFFFFFFF])

    # ------------------------------------------------------------------ graphs
    def capture(self, steps):
        """Record `steps` time steps (from the current step parity on) into one HIP graph; each
        :meth:`replay` of it advances the ensemble by `steps` (even, so the parity stays in step)."""
        if steps % 2:
            raise ValueError("capture an even number of steps")
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        start = self.steps_done
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for t in range(steps):
                    self.sweep(step=start + t)
        torch.cuda.current_stream(self.device).wait_stream(s)
        g.lbm_steps = steps
        return g

    def replay(self, graph):
        """One launch of a captured graph: advances every member by the graph's step count."""
        graph.replay()
        self.steps_done += graph.lbm_steps

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            nat.check(self.lib.lbm355_batch_destroy(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
