"""GPU macroscopic setter / getter (lbmpy macroscopic_value_kernels.py) for every streaming pattern and the
three collision families, against the FP64 oracle: the setter's populations read through the previous
kernel's write accessor equal the oracle's equilibrium (force shift included), setter -> getter returns the
fields, and the getter after sweeps equals the oracle's density and velocity.  LatticeBoltzmannStep with a
central-moment method (whose setter collides with every rate 1) matches the oracle."""
import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method, Timestep
from lbmpy_amd.macroscopic_value_kernels import (compile_macroscopic_values_setter, compile_macroscopic_values_getter,
                                                 macroscopic_values_setter, macroscopic_values_getter)
from lbmpy_amd.sweep import create_pdf_array
from oracle.lbm_oracle import OracleDomain, OracleCollision, equilibrium_pdfs
from tests.lbm_helpers import seeded_state, storage_view
from tests.test_gpu_streaming_patterns import _natural, _gpu_run

pytestmark = pytest.mark.gpu

METHODS = [
    ("D2Q9-TRT-force", dict(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rates=[1.7, 1.1],
                            force=(1e-4, 2e-5)), (12, 9)),
    ("D3Q19-central", dict(stencil=LBStencil(Stencil.D3Q19), method=Method.CENTRAL_MOMENT, relaxation_rate=1.7,
                           compressible=True, force=(1e-5, 0, 2e-5)), (8, 6, 5)),
    ("D3Q27-cumulant", dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.9,
                            compressible=True), (6, 5, 7)),
    ("D3Q27-central-incompressible", dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CENTRAL_MOMENT,
                                          relaxation_rate=1.6), (5, 6, 4)),
]
PATTERNS = ["pull", "push", "aa", "esotwist", "esopull", "esopush"]


def _fields(shape, D, seed):
    rng = np.random.default_rng(seed)
    grid = tuple(reversed(shape))
    rho = 1.0 + 0.01 * rng.standard_normal(grid)
    u = 0.04 * rng.standard_normal((D,) + grid)
    return rho, u


def _to_tensors(rho, u, shape, D, g=1):
    """[x, y, (z)] density and [x, y, (z), D] velocity tensors with a ghost layer (interior set)."""
    inner = (slice(g, -g),) * D if g else (slice(None),) * D
    r = torch.zeros(tuple(n + 2 * g for n in shape), dtype=torch.float64, device="cuda")
    v = torch.zeros(tuple(n + 2 * g for n in shape) + (D,), dtype=torch.float64, device="cuda")
    r[inner] = torch.from_numpy(np.ascontiguousarray(rho.T)).cuda()
    v[inner] = torch.from_numpy(np.ascontiguousarray(np.moveaxis(u, 0, -1).transpose(*reversed(range(D)), D))).cuda()
    return r, v


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("case", METHODS, ids=lambda c: c[0])
def test_setter_matches_oracle_equilibrium_and_getter_round_trips(case, pattern):
    name, kw, shape = case
    m = create_lb_method(LBMConfig(**kw))
    D = m.dim
    rho, u = _fields(shape, D, 5)
    r, v = _to_tensors(rho, u, shape, D)
    prev = Timestep.BOTH if pattern in ("pull", "push") else Timestep.EVEN
    pdfs = create_pdf_array(shape, m.stencil.Q, 1, "float64")
    print(f"setter {name} {pattern}", flush=True)
    setter = compile_macroscopic_values_setter(m, {"density": r, "velocity": v}, streaming_pattern=pattern,
                                               previous_timestep=prev, periodicity=(True,) * D)
    setter(pdfs)
    key = (pattern, 2) if pattern in ("pull", "push") else (pattern, 0)
    arr = storage_view(pdfs, D).cpu().numpy()
    got = _natural(arr, m, shape, key, 1, (True,) * D)
    want = equilibrium_pdfs(m, rho, u)
    assert np.abs(got - want).max() <= 1e-14
    getter = compile_macroscopic_values_getter(m, ["density", "velocity"], streaming_pattern=pattern,
                                               previous_timestep=prev, periodicity=(True,) * D)
    r2, v2 = torch.zeros_like(r), torch.zeros_like(v)
    getter(pdfs, density=r2, velocity=v2)
    inner = (slice(1, -1),) * D
    assert torch.abs(r2[inner] - r[inner]).max().item() <= 1e-14
    assert torch.abs(v2[inner] - v[inner]).max().item() <= 1e-14


@pytest.mark.parametrize("pattern", ["aa", "esotwist", "push"])
def test_getter_after_sweeps_matches_oracle(pattern):
    name, kw, shape = METHODS[0]
    m = create_lb_method(LBMConfig(**kw))
    D = m.dim
    f0 = seeded_state(m, shape, seed=8, du=0.04, noise=1e-3)
    steps = 6
    dom = OracleDomain(m, shape, faces={a: "periodic" for a in range(D)}, pattern=pattern)
    dom.set_pdfs_natural(f0)
    for _ in range(steps):
        dom.step()
    f = dom.pdfs_in_natural_layout()
    rho_o, _, _, u_o = OracleCollision(m, {}).macroscopic(f.reshape(m.stencil.Q, -1))
    # GPU: sweep, then the getter through the last kernel's write accessor
    print(f"getter after sweeps {pattern}", flush=True)
    from tests.test_gpu_streaming_patterns import _kernels, _place
    periodic = (True,) * D
    ks = _kernels(kw, pattern)
    for k in ks:
        k.set_faces(["periodic"] * D)
    first = ("push", 2) if pattern == "push" else (pattern, 0)
    src = create_pdf_array(shape, m.stencil.Q, 1, "float64")
    storage_view(src, D)[...] = torch.from_numpy(_place(f0, m, shape, first, 1, periodic)).cuda()
    if pattern == "push":
        dst = create_pdf_array(shape, m.stencil.Q, 1, "float64")
        for _ in range(steps):
            ks[0](src=src, dst=dst)
            src, dst = dst, src
        prev = Timestep.BOTH
    else:
        even, odd = ks
        for t in range(steps):
            (odd if t % 2 == 0 else even)(src=src)
        prev = Timestep.EVEN
    r = torch.zeros(tuple(n + 2 for n in shape), dtype=torch.float64, device="cuda")
    v = torch.zeros(tuple(n + 2 for n in shape) + (D,), dtype=torch.float64, device="cuda")
    k = macroscopic_values_getter(m, r, v, src, streaming_pattern=pattern, previous_timestep=prev,
                                  periodicity=periodic).compile()
    k()
    inner = (slice(1, -1),) * D
    got_r = r[inner].cpu().numpy().T.reshape(-1)
    got_u = np.moveaxis(v[inner].cpu().numpy(), -1, 0)
    got_u = np.stack([got_u[a].T.reshape(-1) for a in range(D)])
    assert np.abs(got_r - rho_o).max() <= 1e-14
    assert np.abs(got_u - np.asarray(u_o)).max() <= 1e-14


def test_lbstep_central_moment_channel():
    """LatticeBoltzmannStep with a central-moment method (its setter needs the equilibrium-by-collision
    path), D3Q19, force-driven channel with NoSlip walls, against the oracle after 40 steps."""
    from lbmpy_amd.lbstep import LatticeBoltzmannStep
    from lbmpy_amd import CreateKernelConfig, Target
    from lbmpy_amd.boundaries import NoSlip, slice_from_direction
    lc = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.CENTRAL_MOMENT, relaxation_rate=1.8,
                   compressible=True, force=(1e-5, 0, 0))
    step = LatticeBoltzmannStep(domain_size=(10, 8, 6), periodicity=(True, False, True), lbm_config=lc,
                                config=CreateKernelConfig(target=Target.GPU, data_type="float64"))
    for w in ("N", "S"):
        step.boundary_handling.set_boundary(NoSlip(), slice_from_direction(w, 3))
    print("lbstep central", flush=True)
    step.run(40)
    from tests.test_gpu_lbstep import _natural as step_natural
    got = step_natural(step)
    m = step.method
    dom = OracleDomain(m, (10, 8, 6), faces={0: "periodic", 1: "noslip", 2: "periodic"})
    rho = np.ones((6, 8, 10))
    dom.set_pdfs_interior(equilibrium_pdfs(m, rho, np.zeros((3,) + rho.shape)))
    for _ in range(40):
        dom.step()
    assert np.abs(got - dom.pdfs_in_natural_layout()).max() <= 1e-12
    assert np.abs(step.velocity[:, :, :]).max() > 1e-6


_SETGET_METHODS = {}


def _setget_method(stencil, force_model, compressible, zero_centered):
    from lbmpy_amd import ForceModel  # noqa: F401
    key = (stencil, force_model, compressible, zero_centered)
    if key not in _SETGET_METHODS:
        st = LBStencil(stencil)
        _SETGET_METHODS[key] = create_lb_method(LBMConfig(
            stencil=st, force_model=force_model, method=Method.TRT, compressible=compressible,
            zero_centered=zero_centered, force=(0.1, 0.12, -0.17)[:st.D]))
    return _SETGET_METHODS[key]


@pytest.mark.parametrize("prev_timestep", ["EVEN", "ODD"])
@pytest.mark.parametrize("streaming_pattern", ["push", "pull", "aa", "esotwist", "esopull", "esopush"])
@pytest.mark.parametrize("zero_centered", [True, False])
@pytest.mark.parametrize("compressible", [True, False])
@pytest.mark.parametrize("force_model", ["GUO", "LUO", None])
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19"])
def test_set_get_density_velocity_with_fields(stencil, force_model, compressible, zero_centered, streaming_pattern,
                                              prev_timestep):
    """lbmpy's tests/test_macroscopic_value_kernels.py:13-57: random density and velocity fields set through the
    accessor of the kernel that ran last (either parity) come back from the getter, with density_deviation."""
    from lbmpy_amd import ForceModel, Timestep
    from lbmpy_amd.macroscopic_value_kernels import compile_macroscopic_values_getter, \
        compile_macroscopic_values_setter
    m = _setget_method(Stencil[stencil], None if force_model is None else ForceModel[force_model], compressible,
                       zero_centered)
    D = m.dim
    g = 1
    inner_size = (3, 7, 4)[:D]
    total = tuple(s + 2 * g for s in inner_size)
    rng = np.random.default_rng(7)
    inner = (slice(g, -g),) * D
    rho_in = torch.zeros(total, dtype=torch.float64, device="cuda")
    u_in = torch.zeros(total + (D,), dtype=torch.float64, device="cuda")
    rho_in[inner] = torch.from_numpy(1 + 0.2 * (rng.random(inner_size) - 0.5)).cuda()
    u_in[inner] = torch.from_numpy(0.1 * (rng.random(inner_size + (D,)) - 0.5)).cuda()
    pdf_arr = create_pdf_array(inner_size, m.stencil.Q, g, "float64")
    ts = Timestep.EVEN if prev_timestep == "EVEN" else Timestep.ODD
    setter = compile_macroscopic_values_setter(m, {"density": rho_in, "velocity": u_in}, pdf_arr=pdf_arr,
                                               ghost_layers=g, streaming_pattern=streaming_pattern,
                                               previous_timestep=ts)
    setter(pdf_arr)
    getter = compile_macroscopic_values_getter(m, ["density", "density_deviation", "velocity"], pdf_arr=pdf_arr,
                                               ghost_layers=g, streaming_pattern=streaming_pattern,
                                               previous_timestep=ts)
    rho_out, dev_out, u_out = torch.zeros_like(rho_in), torch.zeros_like(rho_in), torch.zeros_like(u_in)
    getter(pdfs=pdf_arr, density=rho_out, density_deviation=dev_out, velocity=u_out)
    np.testing.assert_allclose(rho_out.cpu().numpy(), rho_in.cpu().numpy(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(dev_out[inner].cpu().numpy(), rho_in[inner].cpu().numpy() - 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(u_out.cpu().numpy(), u_in.cpu().numpy(), rtol=0, atol=1e-12)
