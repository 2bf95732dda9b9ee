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


def cpu_baseline(n_cells_side=16, steps=4, budget_s=20.0):
    """The FP64 oracle (numpy, 1 thread) on a bounded sample of the same workload."""
    import numpy as np
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
    from oracle.lbm_oracle import OracleDomain, equilibrium_pdfs
    n = n_cells_side
    omega = 1.9848
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT,
                                   relaxation_rate=omega, compressible=True))
    k = 2 * np.pi / n
    c = (np.arange(n) + 0.5) * k
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    u0 = 0.05
    u = np.stack([u0 * np.sin(x) * np.cos(y) * np.cos(z), -u0 * np.cos(x) * np.sin(y) * np.cos(z), 0 * x])
    rho = 1 + 3 * (u0 ** 2 / 16) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
    dom = OracleDomain(m, (n, n, n), pattern="aa")
    f = equilibrium_pdfs(m, rho, u)
    dom.set_pdfs_interior(f[dom.inv])
    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    if ctx:
        ctx.__enter__()
    try:
        dom.step()
        t0 = time.perf_counter()
        done = 0
        while done < steps or time.perf_counter() - t0 < 1.0:
            dom.step()
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
    finally:
        if ctx:
            ctx.__exit__(None, None, None)
    return {"value": n ** 3 * done / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle, D3Q27 cumulant AA TGV {n}^3, {done} steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--dtype", default="float32", choices=["float32", "float64", "float16"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="after the timed region, copy the PDF field once (a plain copy of known size, "
                         "used to scale the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.steps % 2 or args.warmup % 2:
        print("AA streaming needs even step counts", file=sys.stderr)
        sys.exit(2)

    import torch
    from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target,
                           Timestep)
    from lbmpy_amd.sweep import create_pdf_array

    n = args.n
    u0 = 0.05
    nu = u0 * (n / (2 * math.pi)) / 1600.0
    omega = 1.0 / (3.0 * nu + 0.5)
    compute = "float64" if args.dtype == "float64" else "float32"
    cfg = CreateKernelConfig(target=Target.GPU, data_type=args.dtype, compute_type=compute, ghost_layers=0)
    base = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=omega,
                compressible=True, zero_centered=True, streaming_pattern="aa")
    even = create_lb_function(lbm_config=LBMConfig(**base, timestep=Timestep.EVEN), config=cfg)
    odd = create_lb_function(lbm_config=LBMConfig(**base, timestep=Timestep.ODD), config=cfg)
    for k in (even, odd):
        k.set_faces(["periodic"] * 3)

    pdf = create_pdf_array((n, n, n), 27, ghost_layers=0, dtype=args.dtype)
    storage = pdf.permute(3, 2, 1, 0)
    rho, u = tgv_fields(n, u0)
    feq = d3q27_cumulant_equilibrium(rho, u)
    from lbmpy_amd.stencils import LBStencil as _S
    inv = _S("D3Q27").inverse_indices()
    storage.copy_(feq[inv].to(storage.dtype))            # AA initial layout: slot inv(i) holds f_i
    del feq, rho, u
    torch.cuda.synchronize()

    kernels = (odd, even)
    for t in range(args.warmup):
        kernels[t % 2](src=pdf)
    torch.cuda.synchronize()

    # per-launch timing of both parities (events on the launch stream)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(8)]
    for j, (a, b) in enumerate(ev):
        a.record()
        kernels[j % 2](src=pdf)
        b.record()
    torch.cuda.synchronize()
    per = [a.elapsed_time(b) for a, b in ev]
    odd_ms = sum(per[0::2]) / 4
    even_ms = sum(per[1::2]) / 4

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for t in range(args.steps):
        kernels[t % 2](src=pdf)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    ms_per_step = ms / args.steps
    cells = n ** 3
    mlups = cells * args.steps / (ms / 1e3) / 1e6
    s = {"float64": 8, "float32": 4, "float16": 2}[args.dtype]
    bytes_per_cell = 2 * 27 * s
    achieved = bytes_per_cell * cells / (ms_per_step / 1e3) / 1e9

    finite = bool(torch.isfinite(storage[:, ::16, ::16, ::16]).all().item())
    copy_gbs = None
    if args.calibrate:
        # plain copies of the PDF field (known size) with regular and non-temporal accesses:
        # the counter scale for FETCH_SIZE / WRITE_SIZE and the achievable copy bandwidth
        import ctypes
        from lbmpy_amd import _native as nat
        lib = nat.load()
        other = torch.empty_like(storage)
        nfl = storage.numel() * storage.element_size() // 4
        copy_gbs = {}
        for nt in (0, 1):
            nat.check(lib.lbm355_stream_copy(ctypes.c_void_p(storage.data_ptr()), ctypes.c_void_p(other.data_ptr()),
                                             nfl, nt, nat.current_stream_handle()))
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(5):
                nat.check(lib.lbm355_stream_copy(ctypes.c_void_p(storage.data_ptr()),
                                                 ctypes.c_void_p(other.data_ptr()), nfl, nt,
                                                 nat.current_stream_handle()))
            b.record()
            torch.cuda.synchronize()
            copy_gbs["nontemporal" if nt else "regular"] = round(2 * nfl * 4 * 5 / (a.elapsed_time(b) / 1e3) / 1e9, 1)
        print(f"calibration copy: {nfl * 4} bytes read and written per dispatch", flush=True)
        del other
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "MLUPS & % HBM roofline (D3Q27 cumulant 256^3), per-form err vs FP64 oracle",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"float32": "f32", "float64": "f64", "float16": "f16-storage/f32"}[args.dtype],
        "data": "synthetic (Taylor-Green vortex initial state)",
        "config": {"workload": f"D3Q27 cumulant AA stream-collide, periodic {n}^3 TGV, u0=0.05, Re=1600 "
                               f"(omega={omega:.5f}), zero-centred {args.dtype} PDFs",
                   "cells": cells, "bytes_per_cell_step": bytes_per_cell},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": {"aa_odd": round(odd_ms, 4), "aa_even": round(even_ms, 4)},
                     "copy_gbs": copy_gbs},
        "cpu_baseline": cpu,
        "finite": finite,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
