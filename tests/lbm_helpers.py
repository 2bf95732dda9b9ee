"""Shared helpers for the parity tests: run the same seeded setup through the oracle
(CPU, FP64) and through the HIP sweep (GPU), in lbmpy's time-loop order."""
import numpy as np

from lbmpy_amd import LBMConfig, create_lb_method
from oracle.lbm_oracle import OracleDomain, equilibrium_pdfs


def seeded_state(method, shape, seed=0, du=0.02, drho=0.01, noise=1e-4):
    """Equilibrium of a smooth random (rho, u) field plus a small non-equilibrium part,
    as interior PDFs shaped (Q, [nz,] ny, nx)."""
    rng = np.random.default_rng(seed)
    D = method.dim
    grid = tuple(reversed(shape))
    rho = 1.0 + drho * rng.standard_normal(grid)
    u = du * rng.standard_normal((D,) + grid)
    f = equilibrium_pdfs(method, rho, u)
    f = f + noise * rng.standard_normal(f.shape)
    return f


def oracle_run(method, shape, faces, pattern, f0, steps, bindings=None, collision=None):
    dom = OracleDomain(method, shape, faces=faces, pattern=pattern, bindings=bindings, collision=collision)
    if pattern == "aa":
        dom.set_pdfs_interior(f0[dom.inv])
    else:
        dom.set_pdfs_interior(f0)
    for _ in range(steps):
        dom.step()
    return dom.pdfs_in_natural_layout(), dom


def gpu_kernels(lbm_config_kwargs, data_type="float64", compute_type=None, ghost_layers=1):
    """Sweep kernel(s) for the given LBMConfig keyword arguments on the GPU path."""
    from lbmpy_amd import create_lb_function, CreateKernelConfig, Target, Timestep
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type, compute_type=compute_type,
                             ghost_layers=ghost_layers)
    pattern = lbm_config_kwargs.get("streaming_pattern", "pull")
    if pattern == "aa":
        return [create_lb_function(lbm_config=LBMConfig(**lbm_config_kwargs, timestep=t), config=cfg)
                for t in (Timestep.EVEN, Timestep.ODD)]
    return [create_lb_function(lbm_config=LBMConfig(**lbm_config_kwargs), config=cfg)]


def storage_view(arr, D):
    """(Q, [Z,] Y, X) view of an [x, y, (z,) q] PDF tensor."""
    return arr.permute(*range(D, -1, -1))


def gpu_run(kernels, method, shape, faces, f0, steps, data_type="float64", ghost_layers=1, bindings=None):
    import torch
    from lbmpy_amd.sweep import create_pdf_array
    D = method.dim
    Q = method.stencil.Q
    names = {"periodic": "periodic", "noslip": "noslip", "ghost": "ghost"}
    for k in kernels:
        k.set_faces([names[faces[a]] for a in range(D)])
    aa = len(kernels) == 2
    inv = np.array(method.stencil.inverse_indices())
    g = ghost_layers
    src = create_pdf_array(shape, Q, g, data_type)
    sl = (slice(None),) + (slice(g, -g if g else None),) * D
    init = f0[inv] if aa else f0
    storage_view(src, D)[sl] = torch.from_numpy(init).to(src.dtype).cuda()
    bindings = bindings or {}
    if aa:
        assert steps % 2 == 0, "compare AA runs after an even number of steps (EVEN kernel last)"
        even, odd = kernels
        for t in range(steps):
            (odd if t % 2 == 0 else even)(src=src, **bindings)
        out = storage_view(src, D)[sl].double().cpu().numpy()
        return out[inv] if steps % 2 == 0 else out
    dst = create_pdf_array(shape, Q, g, data_type)
    k = kernels[0]
    for _ in range(steps):
        k(src=src, dst=dst, **bindings)
        src, dst = dst, src
    return storage_view(src, D)[sl].double().cpu().numpy()
