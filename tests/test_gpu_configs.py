"""Two benchmark configurations at their real sizes, and the batched path against the oracle.

* D3Q19 MRT (weighted-orthogonal) Poiseuille channel at 512 x 256 x 256: NoSlip walls in y, periodic x and z,
  Guo force 1e-6, omega_shear 1.8 and every other rate 1, FP32 arithmetic on zero-centred PDFs stored in FP64,
  FP32 and FP16, pull streaming, 2000 steps (the configuration runs 20000; the start-up series is exact at any
  t).  FP32 storage equals FP64 storage bit for bit (every stored value is an FP32 number), the x-velocity
  profile follows the analytic start-up of plane Poiseuille flow (analytical_solutions.py:5-20 for its steady
  limit) to 4e-3 of its maximum (the lattice's own error near the walls at this early time is 2.7e-3), and the
  FP64 mass sum drifts by less than 1e-7 per cell.  FP16 storage with round-to-nearest drops the force (its
  per-step increment is below half an FP16 ulp; the sweep warns) and is 1.4e-2 off the FP32 run already; with
  stochastic rounding it stays within 2e-3 of the FP32 run (7e-4 measured).  A small FP32-arithmetic copy of the
  channel is compared with the FP64 oracle.
* 256 D3Q19 TRT channels of seeded uneven sizes {32..96 step 8}^3 (numpy default_rng(2024)), per-member force
  log-uniform in [1e-7, 1e-5] and omega uniform in [1.5, 1.95], NoSlip y, FP32 pull, 1000 steps as one replay
  of a captured HIP graph of the autotuned batched sweep: every member equals its single-domain run bit for bit.
* An FP64 ensemble (D3Q19 TRT with walls and force, D2Q9 MRT) against the oracle member by member, <= 1e-12.
"""
import numpy as np
import pytest
import torch

from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, create_lb_function, create_lb_method,
                       CreateKernelConfig, Target)
from lbmpy_amd.sweep import create_pdf_array

pytestmark = pytest.mark.gpu


def poiseuille_startup(y, t, force, nu, h, terms=4000):
    """Plane Poiseuille start-up from rest under a body force, walls at y = 0 and y = h:
    u = F y (h - y) / (2 nu) - sum_{n odd} 4 F h^2 / (nu pi^3 n^3) sin(n pi y / h) exp(-n^2 pi^2 nu t / h^2)."""
    u = force * y * (h - y) / (2.0 * nu)
    n = np.arange(1, 2 * terms, 2, dtype=np.float64)[:, None]
    return u - (4.0 * force * h * h / (nu * np.pi ** 3 * n ** 3) * np.sin(n * np.pi * y[None, :] / h)
                * np.exp(-(n ** 2) * np.pi ** 2 * nu * t / (h * h))).sum(axis=0)


def _mrt_channel_kernel(data_type, compute_type="float32", ghost_layers=0, stochastic_rounding=False, **extra):
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type, compute_type=compute_type,
                             ghost_layers=ghost_layers, stochastic_rounding=stochastic_rounding)
    k = create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.MRT,
                                                relaxation_rate=1.8, force=(1e-6, 0.0, 0.0), zero_centered=True,
                                                compressible=False, **extra), config=cfg)
    k.set_faces(["periodic", "noslip", "periodic"])
    return k


def _x_velocity_profile(arr, force):
    """u_x(y) at the mid x / mid z column of an [x, y, z, q] zero-centred incompressible field, FP64."""
    c = torch.tensor(LBStencil(Stencil.D3Q19).stencil_entries, dtype=torch.float64, device=arr.device)
    nx, ny, nz = arr.shape[:3]
    col = arr[nx // 2, :, nz // 2, :].double()                  # (ny, 19)
    return ((col * c[:, 0]).sum(1) - force / 2.0).cpu().numpy()  # u = j / rho_0 + F / 2 with rho_0 = 1


def test_mrt_channel_512x256x256_storage_forms():
    nx, ny, nz = 512, 256, 256
    steps, force, omega = 2000, 1e-6, 1.8
    nu = (1.0 / omega - 0.5) / 3.0
    y = np.arange(ny) + 0.5
    u_an = poiseuille_startup(y, steps - 0.5, force, nu, float(ny))
    final, profiles = {}, {}
    import warnings
    for dt, sr in (("float64", False), ("float32", False), ("float16", False), ("float16", True)):
        print(f"MRT channel 512x256x256 {dt} storage{' (stochastic rounding)' if sr else ''}, {steps} steps",
              flush=True)
        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            k = _mrt_channel_kernel(dt, stochastic_rounding=sr)
        warned = any(issubclass(w.category, RuntimeWarning) for w in caught)
        assert warned == (dt == "float16" and not sr)
        src = create_pdf_array((nx, ny, nz), 19, ghost_layers=0, dtype=dt)
        dst = create_pdf_array((nx, ny, nz), 19, ghost_layers=0, dtype=dt)
        mass0 = k.conserved(src)[0]
        for _ in range(steps):
            k(src=src, dst=dst)
            src, dst = dst, src
        torch.cuda.synchronize()
        mass1 = k.conserved(src)[0]
        prof = _x_velocity_profile(src, force)
        err = np.abs(prof - u_an).max() / np.abs(u_an).max()
        drift = (mass1 - mass0) / (nx * ny * nz)
        print(f"  profile error vs analytic {err:.3e}, mass drift per cell {drift:.3e}", flush=True)
        key = dt + ("-sr" if sr else "")
        profiles[key] = prof
        if dt != "float16":
            assert err <= 4e-3
            assert abs(drift) <= 1e-7
            # the profile is the same in every x-z column of the periodic directions
            assert np.abs(_x_velocity_profile(src[nx // 3:], force) - prof).max() <= 1e-9
            final[dt] = src
        else:
            assert abs(drift) <= 1e-5
        del dst
    assert torch.equal(final["float64"].float(), final["float32"])
    umax = np.abs(u_an).max()
    rne = np.abs(profiles["float16"] - profiles["float32"]).max() / umax
    sr = np.abs(profiles["float16-sr"] - profiles["float32"]).max() / umax
    print(f"FP16 storage vs FP32 storage: round-to-nearest {rne:.3e}, stochastic rounding {sr:.3e}", flush=True)
    assert sr <= 2e-3
    assert rne >= 5e-3


def test_mrt_channel_fp32_arithmetic_against_fp64_oracle():
    """FP32 arithmetic and storage on a 16 x 12 x 6 copy of the channel, 300 steps, against the FP64 oracle:
    populations within 2e-7 and the velocity within 1e-3 of its maximum."""
    from oracle.lbm_oracle import OracleDomain, equilibrium_pdfs
    shape, steps = (16, 12, 6), 300
    k = _mrt_channel_kernel("float32", ghost_layers=1)
    m = k.method
    dom = OracleDomain(m, shape, faces={0: "periodic", 1: "noslip", 2: "periodic"}, pattern="pull")
    grid = tuple(reversed(shape))
    dom.set_pdfs_interior(equilibrium_pdfs(m, np.ones(grid), np.zeros((3,) + grid)))
    src = create_pdf_array(shape, 19, 1, "float32")
    from tests.lbm_helpers import storage_view
    storage_view(src, 3)[(slice(None),) + (slice(1, -1),) * 3] = torch.from_numpy(dom.pdfs_interior()).float().cuda()
    dst = create_pdf_array(shape, 19, 1, "float32")
    for _ in range(steps):
        k(src=src, dst=dst)
        src, dst = dst, src
        dom.step()
    got = storage_view(src, 3)[(slice(None),) + (slice(1, -1),) * 3].double().cpu().numpy()
    want = dom.pdfs_interior()
    assert np.abs(got - want).max() <= 2e-7
    c = np.array(m.stencil.stencil_entries, dtype=np.float64)
    ux_got = np.tensordot(c[:, 0], got, axes=(0, 0))
    ux_want = np.tensordot(c[:, 0], want, axes=(0, 0))
    assert np.abs(ux_want).max() > 1e-5
    assert np.abs(ux_got - ux_want).max() <= 1e-3 * np.abs(ux_want).max()


def test_ensemble_256_members_graph_replay_bit_exact():
    from lbmpy_amd.batch import BatchedSweep
    rng = np.random.default_rng(2024)
    n, steps = 256, 1000
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
    print(f"ensemble: {n} members, {cells} cells, {steps} steps", flush=True)
    b = BatchedSweep(kernels, src, dst)
    tuned = b.autotune(candidates=[(256, 0, 0), (256, 4, 1), (128, 8, 1), (64, 8, 0)], reps=3)
    print(f"ensemble: autotuned launch {tuned['launch']}", flush=True)
    for a in src + dst:
        a.zero_()
    b.steps_done = 0
    graph = b.capture(steps)
    b.replay(graph)
    torch.cuda.synchronize()
    batched = [a.clone() for a in b.current_arrays()]
    del graph
    b.close()
    print("ensemble: single-domain runs", flush=True)
    for a in src + dst:
        a.zero_()
    for k, s_, d_ in zip(kernels, src, dst):
        x, y = s_, d_
        for _ in range(steps):
            k(src=x, dst=y)
            x, y = y, x
    torch.cuda.synchronize()
    singles = src if steps % 2 == 0 else dst
    exact = [bool(torch.equal(p, q)) for p, q in zip(batched, singles)]
    assert all(exact), f"{sum(exact)}/{n} members bit-exact"
    assert all(float(a.abs().max()) > 0 for a in singles)


ORACLE_ENSEMBLE = [
    ("D3Q19", dict(method=Method.TRT, relaxation_rate=1.7, force=(2e-5, 0.0, 0.0)), (12, 8, 6)),
    ("D3Q19", dict(method=Method.TRT, relaxation_rate=1.55, force=(1e-5, 0.0, 0.0)), (8, 10, 7)),
    ("D3Q19", dict(method=Method.TRT, relaxation_rate=1.9, force=(3e-5, 0.0, 0.0)), (10, 6, 9)),
]


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_fp64_ensemble_matches_oracle(pattern):
    from lbmpy_amd import Timestep
    from lbmpy_amd.batch import BatchedSweep
    from oracle.lbm_oracle import OracleDomain
    from tests.lbm_helpers import seeded_state
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float64", ghost_layers=0)
    kernels, src, dst, refs = [], [], [], []
    steps = 12
    faces = {0: "periodic", 1: "noslip", 2: "periodic"}
    for seed, (st, kw, shape) in enumerate(ORACLE_ENSEMBLE):
        lc = dict(stencil=LBStencil(Stencil[st]), streaming_pattern=pattern, compressible=True, **kw)
        if pattern == "aa":
            k = create_lb_function(lbm_config=LBMConfig(**lc, timestep=Timestep.ODD), config=cfg)
        else:
            k = create_lb_function(lbm_config=LBMConfig(**lc), config=cfg)
        k.set_faces(["periodic", "noslip", "periodic"])
        m = create_lb_method(LBMConfig(**lc))
        f0 = seeded_state(m, shape, seed=50 + seed, du=0.03, noise=1e-3)
        dom = OracleDomain(m, shape, faces=faces, pattern=pattern)
        dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
        for _ in range(steps):
            dom.step()
        refs.append((dom.pdfs_in_natural_layout(), m))
        a = create_pdf_array(shape, 19, ghost_layers=0, dtype="float64")
        init = f0[np.array(m.stencil.inverse_indices())] if pattern == "aa" else f0
        a.permute(3, 2, 1, 0)[...] = torch.from_numpy(init).cuda()
        kernels.append(k)
        src.append(a)
        dst.append(torch.empty_like(a))
    b = BatchedSweep(kernels, src, dst if pattern == "pull" else None)
    b.run(steps)
    torch.cuda.synchronize()
    for (want, m), got in zip(refs, b.current_arrays()):
        g = got.permute(3, 2, 1, 0).cpu().numpy()
        if pattern == "aa":
            g = g[np.array(m.stencil.inverse_indices())]         # EVEN kernel last: slot inv(i) holds f_i
        assert np.abs(g - want).max() <= 1e-12
    b.close()
