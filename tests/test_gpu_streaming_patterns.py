"""Push and the in-place Eso patterns (fieldaccess.py:78-87, 160-316) on the GPU: sweeps and
LatticeBoltzmannStep runs against the FP64 oracle's restatement of lbmpy's accessors, max-abs <= 1e-12 on the
post-collision populations read through the last kernel's write accessor; and the FP32 EsoTwist sweep is
bit-identical to the FP32 AA sweep (same per-cell arithmetic, only the addressing differs)."""
import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method, create_lb_function, \
    CreateKernelConfig, Target, Timestep
from lbmpy_amd.boundaries import NoSlip, UBB, slice_from_direction
from lbmpy_amd.sweep import create_pdf_array
from oracle.lbm_oracle import OracleDomain, streaming_accessor
from tests.lbm_helpers import seeded_state, storage_view

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
PATTERNS = ["push", "esotwist", "esopull", "esopush"]
CASES = [
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6, force=(1e-4, 2e-5)), (70, 6)),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.8, 1.2], force=(1e-5, 0, 0)), (10, 6, 5)),
    ("D3Q19", Method.CENTRAL_MOMENT, dict(relaxation_rate=1.7, compressible=True), (8, 5, 6)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.9, compressible=True, force=(1e-5, 0, 0)), (8, 6, 5)),
]


def _index(shape, off, g, periodic):
    """Storage indices (z, y, x) of interior cells shifted by off; periodic axes wrap, others reach the ghost
    layer.  Arrays are shaped [x, y, (z)]."""
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    idx = []
    for a, n in enumerate(shape):
        v = grids[a] + int(off[a])
        if periodic[a]:
            v = v % n
        idx.append(v + g)
    return tuple(reversed(idx))


def _place(f, m, shape, key, g, periodic):
    """Natural populations f (Q, [nz,] ny, nx) into a storage array (Q, [Z,] Y, X) through kernel key's write
    accessor."""
    c, inv = np.array(m.stencil.stencil_entries), np.array(m.stencil.inverse_indices())
    _, _, woff, wslot = streaming_accessor(c, inv, *key)
    out = np.zeros((m.stencil.Q,) + tuple(n + 2 * g for n in reversed(shape)))
    for i in range(m.stencil.Q):
        out[(wslot[i],) + _index(shape, woff[i], g, periodic)] = f[i].T
    return out


def _natural(arr, m, shape, key, g, periodic):
    c, inv = np.array(m.stencil.stencil_entries), np.array(m.stencil.inverse_indices())
    _, _, woff, wslot = streaming_accessor(c, inv, *key)
    return np.stack([arr[(wslot[i],) + _index(shape, woff[i], g, periodic)].T for i in range(m.stencil.Q)])


def _kernels(kw, pattern, data_type="float64", g=1):
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type, ghost_layers=g)
    if pattern in ("pull", "push"):
        return [create_lb_function(lbm_config=LBMConfig(**kw, streaming_pattern=pattern), config=cfg)]
    return [create_lb_function(lbm_config=LBMConfig(**kw, streaming_pattern=pattern, timestep=t), config=cfg)
            for t in (Timestep.EVEN, Timestep.ODD)]


def _gpu_run(kw, m, shape, pattern, f0, steps, g, data_type="float64"):
    D = m.dim
    periodic = (True,) * D
    ks = _kernels(kw, pattern, data_type, g)
    for k in ks:
        k.set_faces(["periodic"] * D)
    Q = m.stencil.Q
    first = ("push", 2) if pattern == "push" else (pattern, 0)
    src = create_pdf_array(shape, Q, g, data_type)
    storage_view(src, D)[...] = torch.from_numpy(_place(f0, m, shape, first, g, periodic)).to(src.dtype).cuda()
    if pattern == "push":
        dst = create_pdf_array(shape, Q, g, data_type)
        for _ in range(steps):
            ks[0](src=src, dst=dst)
            src, dst = dst, src
        last = ("push", 2)
    else:
        even, odd = ks
        for t in range(steps):
            (odd if t % 2 == 0 else even)(src=src)
        last = (pattern, 0 if steps % 2 == 0 else 1)
    arr = storage_view(src, D).double().cpu().numpy()
    return _natural(arr, m, shape, last, g, periodic)


@pytest.mark.parametrize("g", [1, 0])
@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1].name}")
def test_pattern_sweep_fp64_matches_oracle(case, pattern, g):
    st, meth, extra, shape = case
    kw = dict(stencil=LBStencil(Stencil[st]), method=meth, **extra)
    m = create_lb_method(LBMConfig(**kw))
    f0 = seeded_state(m, shape, seed=17, du=0.04, noise=1e-3)
    dom = OracleDomain(m, shape, faces={a: "periodic" for a in range(m.dim)}, pattern=pattern)
    dom.set_pdfs_natural(f0)
    steps = 7
    for _ in range(steps):
        dom.step()
    print(f"pattern {pattern} {st} {meth.name} g={g}", flush=True)
    got = _gpu_run(kw, m, shape, pattern, f0, steps, g)
    assert np.abs(got - dom.pdfs_in_natural_layout()).max() <= FP64_TOL
    assert np.abs(got - f0).max() > 1e-4


def _step(domain, periodicity, lc, data_type="float64"):
    from lbmpy_amd.lbstep import LatticeBoltzmannStep
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type)
    return LatticeBoltzmannStep(domain_size=domain, periodicity=periodicity, lbm_config=lc, config=cfg)


def _step_natural(step):
    D = step.dim
    arr = step.pdf_array.permute(*range(D, -1, -1)).double().cpu().numpy()
    if step._pattern in ("pull", "push"):
        key = (step._pattern, 2)
    else:
        key = (step._pattern, 0 if step._aa_next_odd else 1)
    return _natural(arr, step.method, step.domain_size, key, 1, step.periodicity)


@pytest.mark.parametrize("pattern", PATTERNS)
def test_lbstep_cavity_with_link_walls(pattern):
    """Lid-driven cavity (NoSlip S/E/W, UBB lid N), D2Q9 TRT; with these patterns the walls are boundary links."""
    lc = LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rates=[1.7, 1.1],
                   compressible=True, streaming_pattern=pattern)
    step = _step((14, 10), (False, False), lc)
    for w in ("S", "E", "W"):
        step.boundary_handling.set_boundary(NoSlip(), slice_from_direction(w, 2))
    step.boundary_handling.set_boundary(UBB((0.05, 0.0)), slice_from_direction("N", 2))
    assert step.boundary_handling.faces == ["ghost", "ghost"]
    print(f"lbstep cavity {pattern}", flush=True)
    step.run(31)
    got = _step_natural(step)
    bh = step.boundary_handling
    specs = []
    for k, o in enumerate(bh._objects):
        spec = {"kind": "noslip"} if isinstance(o, NoSlip) else {"kind": "ubb", "velocity": o._velocity,
                                                                 "density": o._density}
        specs.append((bh.flags == k + 1, spec))
    m = step.method
    dom = OracleDomain(m, (14, 10), faces={0: "ghost", 1: "ghost"}, pattern=pattern, boundaries=specs)
    from oracle.lbm_oracle import equilibrium_pdfs
    dom.set_pdfs_natural(equilibrium_pdfs(m, np.ones((10, 14)), np.zeros((2, 10, 14))))
    for _ in range(31):
        dom.step()
    assert np.abs(got - dom.pdfs_in_natural_layout()).max() <= FP64_TOL
    assert np.abs(got).max() > 1e-3


@pytest.mark.parametrize("pattern", ["esopull", "esotwist", "push"])
def test_lbstep_velocity_matches_pull(pattern):
    """The getter after an Eso / push run gives the pull run's velocity (channel with a force, walls y)."""
    res = []
    for p in ("pull", pattern):
        lc = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rates=[1.8, 1.2],
                       force=(1e-5, 0, 0), streaming_pattern=p)
        step = _step((12, 8, 6), (True, False, True), lc)
        for w in ("N", "S"):
            step.boundary_handling.set_boundary(NoSlip(), slice_from_direction(w, 3))
        step.run(20)
        res.append(step.velocity[:, :, :].copy())
    assert np.abs(res[0] - res[1]).max() <= 1e-14
    assert np.abs(res[0]).max() > 1e-5


def test_fp32_esotwist_bit_identical_to_aa():
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.95, compressible=True)
    m = create_lb_method(LBMConfig(**kw))
    shape = (64, 6, 5)
    f0 = seeded_state(m, shape, seed=2, du=0.04, noise=1e-3).astype(np.float32).astype(np.float64)
    print("fp32 esotwist vs aa", flush=True)
    a = _gpu_run(kw, m, shape, "aa", f0, 8, 0, "float32")
    b = _gpu_run(kw, m, shape, "esotwist", f0, 8, 0, "float32")
    np.testing.assert_array_equal(a, b)


ROW_CASES = [
    ("D2Q9", Method.TRT, dict(relaxation_rates=[1.8, 1.1], force=(1e-4, 0)), (128, 6)),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.8, 1.2], force=(1e-5, 0, 0)), (64, 5, 4)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.9, compressible=True), (64, 4, 5)),
]


@pytest.mark.parametrize("g", [1, 0])
@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("case", ROW_CASES, ids=lambda c: c[0])
def test_pattern_row_staged_sweep_fp64_matches_oracle(case, pattern, g):
    """Whole periodic rows (nx % 64 == 0) take the row-staged kernel (x shifts through LDS)."""
    st, meth, extra, shape = case
    kw = dict(stencil=LBStencil(Stencil[st]), method=meth, **extra)
    m = create_lb_method(LBMConfig(**kw))
    f0 = seeded_state(m, shape, seed=23, du=0.04, noise=1e-3)
    dom = OracleDomain(m, shape, faces={a: "periodic" for a in range(m.dim)}, pattern=pattern)
    dom.set_pdfs_natural(f0)
    steps = 6
    for _ in range(steps):
        dom.step()
    print(f"row-staged pattern {pattern} {st} g={g}", flush=True)
    got = _gpu_run(kw, m, shape, pattern, f0, steps, g)
    assert np.abs(got - dom.pdfs_in_natural_layout()).max() <= FP64_TOL
