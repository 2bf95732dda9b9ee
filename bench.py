"""Headline benchmark: D3Q27 cumulant stream-collide sweep, in-place AA streaming, FP32,
fully periodic 256^3 Taylor-Green vortex on one MI355X.

Prints ONE JSON line (MLUPS, HBM roofline fraction of the sweep, CPU baseline).
MLUPS = interior cells x steps / seconds / 1e6 (lbmpy lbstep.py:329-344).
Algorithmic traffic B = 2 Q s bytes per cell per step (Q loads + Q stores of s bytes).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X vendor peak, GB/s


def tgv_fields(n, u0, dtype_np=None):
    """Taylor-Green vortex, cell-centred: u = u0 (sin x cos y cos z, -cos x sin y cos z, 0),
    rho = 1 + 3 p' with p' = u0^2/16 (cos 2x + cos 2y)(cos 2z + 2)."""
    import torch
    k = 2 * math.pi / n
    c = (torch.arange(n, dtype=torch.float64, device="cuda") + 0.5) * k
    z, y, x = torch.meshgrid(c, c, c, indexing="ij")        # arrays indexed [z, y, x]
    ux = u0 * torch.sin(x) * torch.cos(y) * torch.cos(z)
    uy = -u0 * torch.cos(x) * torch.sin(y) * torch.cos(z)
    uz = torch.zeros_like(ux)
    p = (u0 ** 2 / 16.0) * (torch.cos(2 * x) + torch.cos(2 * y)) * (torch.cos(2 * z) + 2.0)
    rho = 1.0 + 3.0 * p
    return rho, (ux, uy, uz)


def d3q27_cumulant_equilibrium(rho, u, zero_centered=True):
    """Equilibrium of the D3Q27 cumulant method (Maxwellian central moments, backward chimera):
    f_i = rho * prod_a phi_{c_ia}(u_a), phi_0 = 2/3 - u^2, phi_(+-1) = (1/3 + u^2 +- u)/2."""
    import torch
    from lbmpy_amd.stencils import LBStencil
    st = LBStencil("D3Q27")

    def phi(c, ua):
        return (2.0 / 3.0 - ua * ua) if c == 0 else 0.5 * (1.0 / 3.0 + ua * ua + c * ua)
    w1 = {0: 2.0 / 3.0, 1: 1.0 / 6.0, -1: 1.0 / 6.0}
    out = torch.empty((27,) + rho.shape, dtype=torch.float64, device=rho.device)
    for i, cvec in enumerate(st):
        f = rho * phi(cvec[0], u[0]) * phi(cvec[1], u[1]) * phi(cvec[2], u[2])
        if zero_centered:
            f = f - w1[cvec[0]] * w1[cvec[1]] * w1[cvec[2]]
        out[i] = f
    return out


def cpu_baseline(n=64, budget_s=12.0):
    """The C restatement of the FP64 oracle (OpenMP over the host cores this process may use, at most 16)
    on a bounded sample of the same workload: the TGV at n^3, AA, for about `budget_s` seconds."""
    import numpy as np
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
    from oracle.c_oracle import COracle
    from oracle.lbm_oracle import equilibrium_pdfs
    threads = min(16, len(os.sched_getaffinity(0)))
    u0 = 0.05
    omega = 1.0 / (3.0 * u0 * (n / (2 * math.pi)) / 1600.0 + 0.5)
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT,
                                   relaxation_rate=omega, compressible=True, streaming_pattern="aa"))
    k = 2 * np.pi / n
    c = (np.arange(n) + 0.5) * k
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    u = np.stack([u0 * np.sin(x) * np.cos(y) * np.cos(z), -u0 * np.cos(x) * np.sin(y) * np.cos(z), 0 * x])
    rho = 1 + 3 * (u0 ** 2 / 16) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
    inv = np.array(m.stencil.inverse_indices())
    f = np.ascontiguousarray(equilibrium_pdfs(m, rho, u)[inv])
    co = COracle(m)
    co.sweep(f, f, "aa_odd", threads)
    co.sweep(f, f, "aa_even", threads)
    t0 = time.perf_counter()
    done = 0
    while done < 2 or time.perf_counter() - t0 < budget_s:
        co.sweep(f, f, "aa_odd", threads)
        co.sweep(f, f, "aa_even", threads)
        done += 2
    dt = time.perf_counter() - t0
    return {"value": round(n ** 3 * done / dt / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"C/OpenMP FP64 oracle (generic moment-matrix + cumulant-recursion restatement), "
                      f"D3Q27 cumulant AA TGV {n}^3, {done} steps in {dt:.1f} s"}


PROFILE_COUNTERS = os.path.join(ROOT, "profiles", "sweep_d3q27_cumulant_aa_256_three_forms_counters.md")


def profiled_traffic(n, dtype):
    """HBM bytes per sweep launch from the committed FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), when it
    covers this workload (256^3, the storage type's line); otherwise None."""
    if n != 256 or not os.path.exists(PROFILE_COUNTERS):
        return None
    import re
    with open(PROFILE_COUNTERS) as fh:
        m = re.search(r"traffic_bytes_per_launch\[" + re.escape(dtype) + r"\]\s*=\s*(\d+)", fh.read())
    return int(m.group(1)) if m else None


def calibration_copies(n, dtype, elem_bytes, reps=5, caps=(0, 4, 3, 2, 1)):
    """Achievable bandwidth of two copies as large as the headline PDF field (27 n^3 elements of `dtype`),
    timed with HIP events on the launch stream after the timed region:
    - flat: one FP32 stream, regular and non-temporal accesses (lbm355_stream_copy); the same copy scales
      the FETCH_SIZE / WRITE_SIZE counters;
    - plane27: the sweep's own access pattern without the collision, 27 aligned planes read and 27 written per
      element index (lbm355_plane_copy), out of place (regular and non-temporal accesses) and in place (as an
      AA sweep accesses memory), each at every occupancy cap in `caps` (waves per SIMD, 0 = none), since the
      copy, like the sweeps, runs fastest below full occupancy.  "plane27*" hold the best cap's figure,
      "plane27_by_waves" all of them.
    GB/s count read + written bytes."""
    import ctypes
    import torch
    from lbmpy_amd import _native as nat
    lib = nat.load()
    code = {"float64": 0, "float32": 1, "float16": 2}[dtype]
    nbytes = 27 * n ** 3 * elem_bytes
    a_buf = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
    b_buf = torch.empty_like(a_buf)
    nfl = a_buf.numel()
    plane = n ** 3
    stream = nat.current_stream_handle()
    pa, pb = ctypes.c_void_p(a_buf.data_ptr()), ctypes.c_void_p(b_buf.data_ptr())

    def timed(fn):
        nat.check(fn())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            nat.check(fn())
        e1.record()
        torch.cuda.synchronize()
        return round(2 * nbytes * reps / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)

    out = {"flat_regular": timed(lambda: lib.lbm355_stream_copy(pa, pb, nfl, 0, stream)),
           "flat_nontemporal": timed(lambda: lib.lbm355_stream_copy(pa, pb, nfl, 1, stream))}
    planes = {"plane27": (pb, 0), "plane27_nontemporal": (pb, 1), "plane27_in_place": (pa, 0)}
    by_waves = {}
    for name, (dst, nt) in planes.items():
        by_waves[name] = {str(w): timed(lambda: lib.lbm355_plane_copy(code, 27, pa, dst, plane, plane, nt, w, stream))
                          for w in caps}
        out[name] = max(by_waves[name].values())
    out["plane27_by_waves"] = by_waves
    print(f"calibration copies: {nbytes} bytes read and written per dispatch: {out}", flush=True)
    del a_buf, b_buf
    torch.cuda.empty_cache()
    return out


def valu_issue_ceiling(reps=5, waves=16384, iters=8192):
    """Achievable VALU issue rate, in billions of wave instructions per second: lbm355_valu_probe runs `waves`
    wavefronts of 8 independent FP32 multiply-add chains (8 * iters VALU instructions each), timed with HIP
    events on the launch stream after the timed region.  The sweeps' SQ_INSTS_VALU per launch divided by their
    launch time is held against it."""
    import ctypes
    import torch
    from lbmpy_amd import _native as nat
    lib = nat.load()
    n = waves * 64
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    stream = nat.current_stream_handle()
    run = lambda: nat.check(lib.lbm355_valu_probe(ctypes.c_void_p(out.data_ptr()), n, iters, stream))
    run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    rate = round(waves * 8.0 * iters * reps / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)
    print(f"VALU issue ceiling: {rate} G wave-instructions/s ({waves} waves x {8 * iters} FMA)", flush=True)
    return rate


def macroscopic_from_aa_storage(storage, zero_centered=True):
    """(rho, u) in FP64 from an AA field after an even number of steps (slot j holds f_inv(j), so the
    velocity sums carry -c_j), evaluated plane by plane to bound the temporary memory."""
    import torch
    from lbmpy_amd.stencils import LBStencil
    st = LBStencil("D3Q27")
    c = torch.tensor([list(d) for d in st], dtype=torch.float64, device=storage.device)
    nz = storage.shape[1]
    rho = torch.empty(storage.shape[1:], dtype=torch.float64, device=storage.device)
    u = torch.empty((3,) + tuple(storage.shape[1:]), dtype=torch.float64, device=storage.device)
    for z0 in range(0, nz, 32):
        g = storage[:, z0:z0 + 32].double()
        r = g.sum(0) + (1.0 if zero_centered else 0.0)
        rho[z0:z0 + 32] = r
        u[:, z0:z0 + 32] = -torch.einsum("qa,q...->a...", c, g) / r
    return rho, u


def aa_schedule(warmup, steps, total_steps):
    """Global step indices of the warm-up, timed and untimed tail launches of an in-place (AA) run.

    One counter runs through all three stretches, so launch t uses kernel t % 2 (0 = ODD, the first AA step,
    1 = EVEN) whatever the split, and the tail runs up to the first even total >= max(total_steps,
    warmup + steps): the AA state is read in the layout an EVEN kernel leaves (slot inv(i) of every cell)."""
    if warmup < 0 or steps < 1 or total_steps < 0:
        raise ValueError("need warmup >= 0, steps >= 1 and total_steps >= 0")
    total = max(total_steps, warmup + steps)
    total += total % 2
    return range(0, warmup), range(warmup, warmup + steps), range(warmup + steps, total)


def run_form(dtype, n, steps, warmup, total_steps, u0=0.05, tune=True):
    """Initialise the TGV, pick the two kernels' occupancy caps by timing them on a copy of the state
    (lbmpy_amd.sweep.tune_occupancy, unless `tune` is false), run `warmup` untimed steps, `steps` timed steps (one
    event pair per launch on the launch stream), then untimed steps up to `total_steps`; returns timings and
    (rho, u)."""
    import torch
    from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target,
                           Timestep)
    from lbmpy_amd.sweep import create_pdf_array

    nu = u0 * (n / (2 * math.pi)) / 1600.0
    omega = 1.0 / (3.0 * nu + 0.5)
    compute = "float64" if dtype == "float64" else "float32"
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype, compute_type=compute, ghost_layers=0)
    base = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=omega,
                compressible=True, zero_centered=True, streaming_pattern="aa")
    even = create_lb_function(lbm_config=LBMConfig(**base, timestep=Timestep.EVEN), config=cfg)
    odd = create_lb_function(lbm_config=LBMConfig(**base, timestep=Timestep.ODD), config=cfg)
    for k in (even, odd):
        k.set_faces(["periodic"] * 3)

    pdf = create_pdf_array((n, n, n), 27, ghost_layers=0, dtype=dtype)
    storage = pdf.permute(3, 2, 1, 0)
    rho, u = tgv_fields(n, u0)
    feq = d3q27_cumulant_equilibrium(rho, u)
    inv = LBStencil("D3Q27").inverse_indices()
    storage.copy_(feq[inv].to(storage.dtype))            # AA initial layout: slot inv(i) holds f_i
    del feq, rho, u
    torch.cuda.synchronize()

    kernels = (odd, even)                                 # the first AA step is ODD
    tuning = None
    if tune:
        from lbmpy_amd.sweep import tune_occupancy
        tuning = tune_occupancy(kernels, pdf, candidates=(0, -1, 2, 3, 4))
    warm, timed, tail = aa_schedule(warmup, steps, total_steps)
    for t in warm:
        kernels[t % 2](src=pdf)
    torch.cuda.synchronize()

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    ev[0].record()
    for i, t in enumerate(timed):
        kernels[t % 2](src=pdf)
        ev[i + 1].record()
    torch.cuda.synchronize()
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    ms = ev[0].elapsed_time(ev[-1])
    odd_ms = [p for p, t in zip(per, timed) if t % 2 == 0]
    even_ms = [p for p, t in zip(per, timed) if t % 2 == 1]

    for t in tail:
        kernels[t % 2](src=pdf)
    torch.cuda.synchronize()
    rho_f, u_f = macroscopic_from_aa_storage(storage)
    finite = bool(torch.isfinite(u_f).all().item())
    del pdf, storage
    torch.cuda.empty_cache()
    avg = (lambda v: sum(v) / len(v) if v else None)
    caps = {"aa_odd": odd.waves_per_simd, "aa_even": even.waves_per_simd,
            "tuning_ms": None if tuning is None else
            {name: {str(c): round(t, 4) for c, t in tuning[i].items()} for i, name in enumerate(("aa_odd", "aa_even"))}}
    return dict(ms=ms, odd_ms=avg(odd_ms), even_ms=avg(even_ms), launch_ms=avg(per), omega=omega, rho=rho_f,
                u=u_f, finite=finite, total_steps=tail.stop, caps=caps)


def poiseuille_startup(y, t, force, nu, h, terms=4000):
    """Start-up of plane Poiseuille flow from rest under a body force (walls at y = 0 and y = h):
    u = F y (h - y) / (2 nu) - sum_{n odd} 4 F h^2 / (nu pi^3 n^3) sin(n pi y / h) exp(-n^2 pi^2 nu t / h^2)."""
    import numpy as np
    u = force * y * (h - y) / (2.0 * nu)
    n = np.arange(1, 2 * terms, 2, dtype=np.float64)[:, None]
    u = u - (4.0 * force * h * h / (nu * np.pi ** 3 * n ** 3) * np.sin(n * np.pi * y[None, :] / h)
             * np.exp(-(n ** 2) * np.pi ** 2 * nu * t / (h * h))).sum(axis=0)
    return u


def run_channel(args):
    """D3Q19 MRT (weighted-orthogonal) force-driven channel: NoSlip walls in y, periodic x and z, Guo force
    1e-6 in x, shear rate 1.8 and every other rate 1, FP32 arithmetic on zero-centred PDFs stored in FP64,
    FP32 and FP16 (round-to-nearest and stochastic rounding); pull streaming.  Reports MLUPS, the velocity
    profile's error against the analytic start-up solution and against the FP64-storage run, and the mass drift
    summed in FP64."""
    import warnings
    import numpy as np
    import torch
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target
    from lbmpy_amd.sweep import create_pdf_array
    nx, ny, nz = args.channel_size
    steps = args.channel_steps
    force, omega = 1e-6, 1.8
    nu = (1.0 / omega - 0.5) / 3.0
    c = torch.tensor(LBStencil(Stencil.D3Q19).stencil_entries, dtype=torch.float64, device="cuda")
    y = np.arange(ny) + 0.5
    u_an = poiseuille_startup(y, steps - 0.5, force, nu, float(ny))
    forms, ref_profile = {}, None
    for dt, sr in (("float64", False), ("float32", False), ("float16", False), ("float16", True)):
        cfg = CreateKernelConfig(target=Target.GPU, data_type=dt, compute_type="float32", ghost_layers=0,
                                 stochastic_rounding=sr, rounding_seed=2024)
        with warnings.catch_warnings():
            # FP16 with round-to-nearest drops this force, which is what the form is run to show
            warnings.simplefilter("ignore", RuntimeWarning)
            k = create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.MRT,
                                                        relaxation_rate=omega, force=(force, 0.0, 0.0),
                                                        zero_centered=True, compressible=False), config=cfg)
        k.set_faces(["periodic", "noslip", "periodic"])
        src = create_pdf_array((nx, ny, nz), 19, ghost_layers=0, dtype=dt)
        dst = create_pdf_array((nx, ny, nz), 19, ghost_layers=0, dtype=dt)
        mass0 = k.conserved(src)[0]
        if not args.no_tune:
            from lbmpy_amd.sweep import tune_occupancy
            tune_occupancy([k], src, dst, candidates=(0, -1, 2, 3, 4))
        for _ in range(min(args.warmup, steps)):
            k(src=src, dst=dst)
            src, dst = dst, src
        done = min(args.warmup, steps)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        timed = min(args.steps, steps - done)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(timed):
            k(src=src, dst=dst)
            src, dst = dst, src
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        for _ in range(steps - done - timed):
            k(src=src, dst=dst)
            src, dst = dst, src
        torch.cuda.synchronize()
        mass1 = k.conserved(src)[0]
        col = src.permute(3, 2, 1, 0)[:, nz // 2, :, nx // 2].double()          # (19, ny)
        rho = 1.0 + col.sum(0)
        ux = ((c[:, 0:1] * col).sum(0) - force / 2.0) / rho
        prof = ux.cpu().numpy()
        umax = float(np.abs(u_an).max())
        res = {"mlups": round(nx * ny * nz * timed / (ms / 1e3) / 1e6, 1),
               "achieved_gbs": round(2 * 19 * src.element_size() * nx * ny * nz * timed / (ms / 1e3) / 1e9, 1),
               "profile_err_vs_analytic_max_rel": float(np.abs(prof - u_an).max() / umax),
               "mass_drift_rel": float((mass1 - mass0) / (nx * ny * nz)),
               "u_centre": float(prof[ny // 2]), "u_centre_analytic": float(u_an[ny // 2]),
               "waves_per_simd": k.waves_per_simd}
        if dt == "float64":
            ref_profile = prof
        else:
            res["profile_err_vs_f64_storage_max_rel"] = float(np.abs(prof - ref_profile).max() / umax)
        name = {"float64": "f64-storage/f32", "float32": "f32", "float16": "f16-storage/f32"}[dt]
        forms[name + (" stochastic-rounding" if sr else "")] = res
        del src, dst
        torch.cuda.empty_cache()
    line = {"metric": "D3Q19 MRT channel: MLUPS, profile error vs analytic start-up and vs FP64 storage, mass drift",
            "workload": f"D3Q19 MRT channel {nx}x{ny}x{nz}, NoSlip y, periodic x/z, Guo F=1e-6, omega_shear=1.8, "
                        f"{steps} steps, pull, zero-centred, FP32 arithmetic",
            "unit": "MLUPS", "forms": forms}
    print(json.dumps(line), flush=True)


def run_ensemble(args):
    """Batched ensemble: 256 D3Q19 TRT channels (NoSlip y, periodic x/z), FP32, pull.  Each axis of a member is
    drawn from {32, 40, ..., 96}, its body force log-uniformly from [1e-7, 1e-5] and its shear rate uniformly
    from [1.5, 1.95] (odd rate by the TRT magic number 3/16), all from numpy's default_rng(seed).  The whole run
    (`ensemble_steps` steps) is one replay of a captured HIP graph of the autotuned batched sweep; the same
    members are then stepped one kernel launch each (lbmpy's per-domain loop) and compared bit for bit."""
    import numpy as np
    import torch
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target
    from lbmpy_amd.batch import BatchedSweep
    from lbmpy_amd.sweep import create_pdf_array
    rng = np.random.default_rng(args.seed)
    n, steps = args.members, args.ensemble_steps
    sizes = np.arange(32, 97, 8)
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float32", ghost_layers=0)
    kernels, shapes = [], []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.choice(sizes, 3))
        force = float(np.exp(rng.uniform(np.log(1e-7), np.log(1e-5))))
        omega = float(rng.uniform(1.5, 1.95))
        k = create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT,
                                                    relaxation_rate=omega, force=(force, 0.0, 0.0)), config=cfg)
        k.set_faces(["periodic", "noslip", "periodic"])
        kernels.append(k)
        shapes.append(shape)
    src = [create_pdf_array(sh, 19, ghost_layers=0, dtype="float32") for sh in shapes]
    dst = [create_pdf_array(sh, 19, ghost_layers=0, dtype="float32") for sh in shapes]
    cells = sum(int(np.prod(sh)) for sh in shapes)
    print(f"ensemble: {n} members, {cells} cells", flush=True)
    b = BatchedSweep(kernels, src, dst)
    tuned = b.autotune()
    print(f"ensemble: autotuned launch {tuned['launch']}", flush=True)
    for a in src + dst:
        a.zero_()
    b.steps_done = 0
    graph = b.capture(steps)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.replay(graph)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    batched = [a.clone() for a in b.current_arrays()]
    del graph
    print(f"ensemble: batched run {ms:.1f} ms", flush=True)
    # lbmpy's way: one launch per member and step
    for a in src + dst:
        a.zero_()
    torch.cuda.synchronize()
    e0.record()
    for k, s_, d_ in zip(kernels, src, dst):
        x, y = s_, d_
        for _ in range(steps):
            k(src=x, dst=y)
            x, y = y, x
    e1.record()
    torch.cuda.synchronize()
    ms_single = e0.elapsed_time(e1)
    singles = src if steps % 2 == 0 else dst
    exact = sum(int(torch.equal(p, q)) for p, q in zip(batched, singles))
    line = {"metric": "batched ensemble MLUPS (one HIP-graph launch) and bit-exactness vs single-domain runs",
            "workload": f"{n} D3Q19 TRT channels, sizes {{32..96 step 8}}^3 (seed {args.seed}), F log-U[1e-7,1e-5], "
                        f"omega U[1.5,1.95], NoSlip y, periodic x/z, FP32 pull, {steps} steps",
            "value": round(cells * steps / (ms / 1e3) / 1e6, 1), "unit": "MLUPS", "cells": cells,
            "ms_total": round(ms, 2), "achieved_gbs": round(152 * cells * steps / (ms / 1e3) / 1e9, 1),
            "single_domain_loop_mlups": round(cells * steps / (ms_single / 1e3) / 1e6, 1),
            "members_bit_exact": f"{exact}/{n}", "autotuned_launch": {"threads": tuned["launch"][0],
                                                                      "blocks_per_cu": tuned["launch"][1],
                                                                      "xcd_order": tuned["launch"][2]},
            "autotune_ms_per_step": tuned["ms"]}
    b.close()
    print(json.dumps(line), flush=True)


def run_closure(args):
    """Learned closure on decaying homogeneous turbulence: D3Q27 cumulant, FP32 PDFs, AA, periodic; coarse
    `closure_coarse`^3 grid, fine grid 4x finer (box filter by 4), seeded solenoidal field with
    E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 0.04, omega_shear = 1.99.  Trained on the fine run of seed 0,
    judged on seed 1 against the filtered fine run, beside Smagorinsky (Cs = 0.12) and no subgrid model."""
    from lbmpy_amd.closure import turbulence_study
    r = turbulence_study(n_coarse=args.closure_coarse, factor=4, steps=args.closure_steps,
                         snap_every=args.closure_steps // 12, train_from=args.closure_steps // 6, epochs=20,
                         batch=65536, seed=0, log=lambda s: print(s, flush=True))
    r.pop("closure")
    models = {k: {"ke_rel_l2_err": round(v["ke_rel_l2_err"], 5), "spectrum_log_err": round(v["spectrum_log_err"], 5),
                  "coarse_mlups": round(v["mlups"], 1)} for k, v in r["models"].items()}
    line = {"metric": "learned closure: kinetic-energy decay and spectrum error vs the box-filtered fine run",
            "workload": f"D3Q27 cumulant decaying turbulence, coarse {r['n_coarse']}^3 (omega 1.99), fine "
                        f"{r['n_fine']}^3 (omega {r['omega_fine']:.5f}), k0=4, u_rms=0.04, {r['steps']} coarse steps, "
                        f"FP32 AA; trained on seed 0, judged on seed 1",
            "models": models, "alpha": r["alpha"], "beta": r["beta"], "offset_only_beta": r["offset_only_beta"],
            "train_loss_first_last": [r["train_loss"][0], r["train_loss"][-1]],
            "train_seconds": round(r["train_seconds"], 2), "fine_mlups": round(r["fine_mlups"], 1),
            "ref_ke": r["ref_ke"], "ke": {k: v["ke"] for k, v in r["models"].items()}}
    print(json.dumps(line), flush=True)


def run_patterns(args):
    """The D3Q27 cumulant sweep (FP32, periodic, no ghost layers) with every streaming pattern -- pull and push
    on two arrays, AA, EsoTwist, EsoPull and EsoPush in place -- with the Galilean / fourth-order corrections, the
    per-rate kernel and the learned closure (AA): MLUPS and achieved bandwidth (2 x 27 x 4 B per cell) per kernel, from HIP events around
    each launch on the launch stream.  The state is the rest equilibrium (timing only; parity is in the tests)."""
    import torch
    from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target,
                           Timestep)
    from lbmpy_amd.sweep import create_pdf_array
    n = args.n
    cells = n ** 3
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float32", ghost_layers=0)
    base = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.99, compressible=True)
    from lbmpy_amd.closure import ClosureModel
    variants = [("pull", "pull", {}), ("push", "push", {}), ("aa", "aa", {}), ("esotwist", "esotwist", {}),
                ("esopull", "esopull", {}), ("esopush", "esopush", {}),
                ("aa+galilean", "aa", dict(galilean_correction=True)),
                ("aa+fourth_order", "aa", dict(fourth_order_correction=0.01)),
                ("aa+galilean+fourth_order", "aa", dict(galilean_correction=True, fourth_order_correction=0.01)),
                # the per-rate kernel (orders >= 3 computed and relaxed with their table rates of 1)
                ("aa per-rate kernel", "aa", {}),
                # the learned closure in place of the shear rate (a seeded network: timing only)
                ("aa+learned_closure", "aa", {})]
    out = {}
    for name, pattern, extra in variants:
        if pattern in ("pull", "push"):
            ks = [create_lb_function(lbm_config=LBMConfig(**base, **extra, streaming_pattern=pattern), config=cfg)]
        else:
            ks = [create_lb_function(lbm_config=LBMConfig(**base, **extra, streaming_pattern=pattern, timestep=t),
                                     config=cfg) for t in (Timestep.ODD, Timestep.EVEN)]
        for k in ks:
            k.set_faces(["periodic"] * 3)
            if name == "aa per-rate kernel":
                k.shortcuts = False
            if name == "aa+learned_closure":
                k.set_closure(ClosureModel(seed=0, nu_scale=2e-3))
        a = create_pdf_array((n, n, n), 27, ghost_layers=0, dtype="float32")
        b = create_pdf_array((n, n, n), 27, ghost_layers=0, dtype="float32") if len(ks) == 1 else a

        def launch(t):
            if len(ks) == 1:
                ks[0](src=a if t % 2 == 0 else b, dst=b if t % 2 == 0 else a)
            else:
                ks[t % 2](src=a)

        warm, timed, _ = aa_schedule(args.warmup, args.steps, 0)
        for t in warm:
            launch(t)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        torch.cuda.synchronize()
        ev[0].record()
        for i, t in enumerate(timed):
            launch(t)
            ev[i + 1].record()
        torch.cuda.synchronize()
        per = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
        launch_ms = sum(per) / len(per)
        gbs = 2 * 27 * 4 * cells / (launch_ms / 1e3) / 1e9
        out[name] = {"mlups": round(cells / (launch_ms / 1e3) / 1e6, 1), "launch_ms": round(launch_ms, 4),
                     "achieved_gbs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
        print(f"{name}: {out[name]}", flush=True)
        del a, b, ks
        torch.cuda.empty_cache()
    line = {"metric": "D3Q27 cumulant sweep MLUPS per streaming pattern and correction (FP32, 256^3 periodic)",
            "workload": f"D3Q27 cumulant, omega=1.99, periodic {n}^3, FP32 zero-centred, rest state, "
                        f"{args.steps} timed launches after {args.warmup} warm-up launches",
            "unit": "MLUPS", "bytes_per_cell_step": 216, "patterns": out}
    print(json.dumps(line), flush=True)


def run_families(args):
    """Every collision family at 256^3 (FP32 storage and arithmetic, AA streaming, periodic, rest state): D3Q19 and
    D3Q27 SRT, TRT and MRT, D3Q27 central-moment and cumulant, each given by its shear rate (lbmpy's defaults for the
    other rates).  Each family's two kernels are tuned for occupancy first (unless --no-tune).  MLUPS and bandwidth
    per family from HIP events around each launch on the launch stream."""
    import torch
    from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target,
                           Timestep)
    from lbmpy_amd.sweep import create_pdf_array, tune_occupancy
    n = args.n
    cells = n ** 3
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float32", ghost_layers=0)
    cases = [("D3Q19 SRT", Stencil.D3Q19, dict(method=Method.SRT, relaxation_rate=1.8)),
             ("D3Q19 TRT", Stencil.D3Q19, dict(method=Method.TRT, relaxation_rates=[1.8, 1.1])),
             ("D3Q19 MRT", Stencil.D3Q19, dict(method=Method.MRT, relaxation_rate=1.8)),
             ("D3Q27 SRT", Stencil.D3Q27, dict(method=Method.SRT, relaxation_rate=1.8)),
             ("D3Q27 TRT", Stencil.D3Q27, dict(method=Method.TRT, relaxation_rates=[1.8, 1.1])),
             ("D3Q27 MRT", Stencil.D3Q27, dict(method=Method.MRT, relaxation_rate=1.8)),
             ("D3Q19 central moment", Stencil.D3Q19, dict(method=Method.CENTRAL_MOMENT, relaxation_rate=1.8,
                                                          compressible=True)),
             ("D3Q27 central moment", Stencil.D3Q27, dict(method=Method.CENTRAL_MOMENT, relaxation_rate=1.8,
                                                          compressible=True)),
             ("D3Q27 central moment, incompressible", Stencil.D3Q27,
              dict(method=Method.CENTRAL_MOMENT, relaxation_rate=1.8, compressible=False)),
             ("D3Q27 cumulant", Stencil.D3Q27, dict(method=Method.CUMULANT, relaxation_rate=1.8, compressible=True))]
    out = {}
    for name, st, kw in cases:
        stencil = LBStencil(st)
        ks = [create_lb_function(lbm_config=LBMConfig(stencil=stencil, streaming_pattern="aa", timestep=t, **kw),
                                 config=cfg) for t in (Timestep.ODD, Timestep.EVEN)]
        for k in ks:
            k.set_faces(["periodic"] * 3)
        a = create_pdf_array((n, n, n), stencil.Q, ghost_layers=0, dtype="float32")
        if not args.no_tune:
            tune_occupancy(ks, a, candidates=(0, -1, 2, 3, 4))
        warm, timed, _ = aa_schedule(args.warmup, args.steps, 0)
        for t in warm:
            ks[t % 2](src=a)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        torch.cuda.synchronize()
        ev[0].record()
        for i, t in enumerate(timed):
            ks[t % 2](src=a)
            ev[i + 1].record()
        torch.cuda.synchronize()
        per = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
        launch_ms = sum(per) / len(per)
        b = 2 * stencil.Q * 4
        gbs = b * cells / (launch_ms / 1e3) / 1e9
        out[name] = {"mlups": round(cells / (launch_ms / 1e3) / 1e6, 1), "launch_ms": round(launch_ms, 4),
                     "bytes_per_cell": b, "achieved_gbs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "waves_per_simd": {"aa_odd": ks[0].waves_per_simd, "aa_even": ks[1].waves_per_simd}}
        print(f"{name}: {out[name]}", flush=True)
        del a, ks
        torch.cuda.empty_cache()
    line = {"metric": "stream-collide MLUPS per collision family (FP32, AA, 256^3 periodic)",
            "workload": f"periodic {n}^3, FP32 zero-centred, rest state, {args.steps} timed launches after "
                        f"{args.warmup} warm-up launches; shear rate 1.8, lbmpy's defaults for the other rates",
            "unit": "MLUPS", "families": out}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--dtype", default="float32", choices=["float32", "float64", "float16"],
                    help="storage type of the headline form")
    ap.add_argument("--total-steps", type=int, default=2000,
                    help="steps each form runs before its velocity/density error vs the FP64 run is taken")
    ap.add_argument("--forms", default="float64,float16",
                    help="other storage forms run beside the headline one ('' for none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's occupancy caps instead of timing the headline kernels at several caps "
                         "(on a copy of the state, before the warm-up) and keeping the fastest")
    ap.add_argument("--workload", default="tgv",
                    choices=["tgv", "channel", "ensemble", "closure", "patterns", "families"],
                    help="tgv: the headline D3Q27 cumulant TGV; channel: D3Q19 MRT channel in three storage forms; "
                         "ensemble: 256 batched D3Q19 TRT channels; closure: learned closure on decaying turbulence; "
                         "patterns: the D3Q27 cumulant sweep with every streaming pattern and correction; "
                         "families: every collision family at 256^3 (FP32, AA)")
    ap.add_argument("--closure-coarse", type=int, default=128)
    ap.add_argument("--closure-steps", type=int, default=1200)
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--channel-size", type=int, nargs=3, default=[512, 256, 256])
    ap.add_argument("--channel-steps", type=int, default=20000)
    ap.add_argument("--no-calibrate", action="store_true",
                    help="skip the calibrations after the timed region (a flat copy of the PDF field's size, "
                         "which also scales the FETCH_SIZE / WRITE_SIZE counters, the 27-plane copy of the "
                         "sweep's access pattern, and the VALU issue probe)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("need --steps >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    if args.workload == "channel":
        run_channel(args)
        return
    if args.workload == "ensemble":
        run_ensemble(args)
        return
    if args.workload == "closure":
        run_closure(args)
        return
    if args.workload == "patterns":
        run_patterns(args)
        return
    if args.workload == "families":
        run_families(args)
        return

    import torch
    n = args.n
    cells = n ** 3
    total = aa_schedule(args.warmup, args.steps, args.total_steps)[2].stop
    forms = [args.dtype] + [f for f in args.forms.split(",") if f and f != args.dtype]
    if "float64" in forms:
        forms = ["float64"] + [f for f in forms if f != "float64"]
    size = {"float64": 8, "float32": 4, "float16": 2}
    tag = {"float32": "f32", "float64": "f64", "float16": "f16-storage/f32"}
    results = {}
    ref = None
    for dt in forms:
        r = run_form(dt, n, args.steps, args.warmup, total, tune=not args.no_tune)
        b = 2 * 27 * size[dt]
        res = {"mlups": round(cells * args.steps / (r["ms"] / 1e3) / 1e6, 1),
               "ms_per_step": round(r["ms"] / args.steps, 4),
               "kernel_ms": {"aa_" + k: (None if r[k + "_ms"] is None else round(r[k + "_ms"], 4))
                             for k in ("odd", "even")},
               "achieved_gbs": round(b * cells / (r["launch_ms"] / 1e3) / 1e9, 1), "bytes_per_cell_step": b,
               "finite": r["finite"], "waves_per_simd": r["caps"]}
        res["roofline_frac"] = round(res["achieved_gbs"] / HBM_PEAK_GBS, 4)
        if dt == "float64":
            ref = (r["rho"], r["u"])
        elif ref is not None:
            du = (r["u"] - ref[1])
            res[f"u_err_vs_f64_after_{total}_steps"] = {
                "max_rel_u0": float(du.abs().max() / 0.05),
                "l2_rel": float(du.norm() / ref[1].norm())}
            res[f"rho_err_vs_f64_after_{total}_steps"] = {"max_abs": float((r["rho"] - ref[0]).abs().max())}
        res["_launch_ms"], res["_ms"], res["_omega"] = r["launch_ms"], r["ms"], r["omega"]
        results[dt] = res
        del r
        torch.cuda.empty_cache()

    head = results[args.dtype]
    copy_gbs = None if args.no_calibrate else calibration_copies(n, args.dtype, size[args.dtype])
    valu_ceiling = None if args.no_calibrate else valu_issue_ceiling()
    cpu = None if args.no_cpu else cpu_baseline()
    b = head["bytes_per_cell_step"]
    achieved = b * cells / (head["_launch_ms"] / 1e3) / 1e9
    traffic = profiled_traffic(n, args.dtype)
    # the fastest 27-plane copy of any access kind and occupancy cap
    best_copy = max(copy_gbs[k] for k in ("plane27", "plane27_nontemporal", "plane27_in_place")) if copy_gbs else None
    forms_out = {tag[k]: {kk: vv for kk, vv in v.items() if not kk.startswith("_")} for k, v in results.items()}
    line = {
        "metric": "MLUPS & % HBM roofline (D3Q27 cumulant 256^3), per-form err vs FP64 oracle",
        "value": head["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": tag[args.dtype],
        "data": "synthetic (Taylor-Green vortex initial state)",
        "config": {"workload": f"D3Q27 cumulant AA stream-collide, periodic {n}^3 TGV, u0=0.05, Re=1600 "
                               f"(omega={head['_omega']:.5f}), zero-centred {args.dtype} PDFs",
                   "cells": cells, "bytes_per_cell_step": b},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "lbm355::sweep_kernel<D3Q27, cumulant, AA odd/even> (both parities, same bytes)",
                     "launch_ms": round(head["_launch_ms"], 4), "kernel_ms": head["kernel_ms"],
                     "algorithmic_bytes_per_launch": b * cells, "copy_gbs": copy_gbs,
                     "frac_of_plane27_copy": round(achieved / best_copy, 4) if best_copy else None,
                     "valu_issue_ceiling_ginst_per_s": valu_ceiling},
        "forms": forms_out,
        "cpu_baseline": cpu,
        "finite": all(v["finite"] for v in results.values()),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
