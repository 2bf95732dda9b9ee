"""Batched ensemble sweep: every member of a batch equals its single-domain run bit for bit, for direct
launches, autotuned launch shapes and a replayed HIP graph; pull and AA; D3Q19 TRT channels (NoSlip y,
periodic x/z) of uneven sizes with per-member force and relaxation rate; and every collision family and
storage type (D3Q27 cumulant, D3Q19 central moments, FP64, FP16 storage)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ensemble(n, seed, pattern, dtype="float32", sizes=(8, 16, 24, 32, 40)):
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target
    from lbmpy_amd import Timestep
    rng = np.random.default_rng(seed)
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype, ghost_layers=0)
    members = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.choice(sizes, 3))
        force = float(np.exp(rng.uniform(np.log(1e-7), np.log(1e-5))))
        omega = float(rng.uniform(1.5, 1.95))
        kw = dict(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rate=omega,
                  force=(force, 0.0, 0.0), streaming_pattern=pattern)
        if pattern == "aa":
            ks = [create_lb_function(lbm_config=LBMConfig(**kw, timestep=t), config=cfg)
                  for t in (Timestep.ODD, Timestep.EVEN)]
        else:
            ks = [create_lb_function(lbm_config=LBMConfig(**kw), config=cfg)]
        for k in ks:
            k.set_faces(["periodic", "noslip", "periodic"])
        members.append((shape, ks))
    return members


def _arrays(members, dtype, seed):
    import torch
    from lbmpy_amd.sweep import create_pdf_array
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = []
    for shape, _ in members:
        a = create_pdf_array(shape, 19, ghost_layers=0, dtype=dtype)
        a.copy_((1e-3 * torch.randn(a.shape, generator=g, device="cuda", dtype=torch.float32)).to(a.dtype))
        out.append(a)
    return out


def _single(members, arrays, steps, pattern):
    import torch
    res = []
    for (shape, ks), a in zip(members, arrays):
        src = a.clone()
        if pattern == "pull":
            dst = torch.empty_like(src)
            for _ in range(steps):
                ks[0](src=src, dst=dst)
                src, dst = dst, src
        else:
            for t in range(steps):
                ks[t % 2](src=src)
        res.append(src)
    return res


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_batch_equals_single_domain_runs(pattern):
    import torch
    from lbmpy_amd.batch import BatchedSweep
    print(f"batch vs single {pattern}", flush=True)
    members = _ensemble(10, seed=5, pattern=pattern)
    init = _arrays(members, "float32", seed=6)
    steps = 20
    ref = _single(members, init, steps, pattern)
    src = [a.clone() for a in init]
    dst = [torch.empty_like(a) for a in init] if pattern == "pull" else None
    kernels = [ks[0] for _, ks in members]
    if pattern == "aa":
        kernels = [ks[0] for _, ks in members]        # ODD kernels: the batch picks the parity itself
    b = BatchedSweep(kernels, src, dst)
    for launch in [(256, 0, 0), (64, 2, 1), (128, 4, 0)]:
        for a, i in zip(src, init):
            a.copy_(i)
        b.steps_done = 0
        b.launch = launch
        b.run(steps)
        torch.cuda.synchronize()
        for k, (got, want) in enumerate(zip(b.current_arrays(), ref)):
            assert torch.equal(got, want), f"member {k} differs with launch {launch}"
    b.close()


def test_batch_graph_replay_and_autotune():
    import torch
    from lbmpy_amd.batch import BatchedSweep
    print("batch graph replay", flush=True)
    members = _ensemble(12, seed=7, pattern="pull", sizes=(16, 24, 32, 40, 48))
    init = _arrays(members, "float32", seed=8)
    ref10 = _single(members, init, 10, "pull")
    ref30 = _single(members, init, 30, "pull")
    src = [a.clone() for a in init]
    dst = [torch.empty_like(a) for a in init]
    b = BatchedSweep([ks[0] for _, ks in members], src, dst)
    tuned = b.autotune(candidates=[(64, 4, 0), (256, 0, 0), (128, 8, 1)], reps=4)
    assert tuple(tuned["launch"]) in [(64, 4, 0), (256, 0, 0), (128, 8, 1)]
    for a, i in zip(src, init):
        a.copy_(i)
    b.steps_done = 0
    g = b.capture(10)
    # capturing does not run anything: the arrays still hold the initial state
    torch.cuda.synchronize()
    assert all(torch.equal(a, i) for a, i in zip(b.current_arrays(), init))
    b.replay(g)
    torch.cuda.synchronize()
    assert all(torch.equal(a, r) for a, r in zip(b.current_arrays(), ref10))
    # eager work between replays, then two more replays
    _ = torch.ones(1000, device="cuda").sum().item()
    b.replay(g)
    torch.cuda.synchronize()
    b.replay(g)
    torch.cuda.synchronize()
    assert b.steps_done == 30
    assert all(torch.equal(a, r) for a, r in zip(b.current_arrays(), ref30))
    b.close()


FAMILY_CASES = [
    # (stencil, method, extra LBMConfig keywords, faces, storage dtype, pattern)
    ("D3Q27", "CUMULANT", dict(compressible=True), ["periodic"] * 3, "float32", "aa"),
    ("D3Q27", "CUMULANT", dict(compressible=True, force=(2e-6, 0.0, 0.0)), ["periodic", "noslip", "periodic"],
     "float32", "pull"),
    ("D3Q19", "CENTRAL_MOMENT", dict(compressible=True), ["periodic", "noslip", "periodic"], "float32", "pull"),
    ("D3Q19", "TRT", dict(force=(1e-6, 0.0, 0.0)), ["periodic", "noslip", "periodic"], "float64", "pull"),
    ("D3Q19", "TRT", dict(force=(1e-6, 0.0, 0.0)), ["periodic", "noslip", "periodic"], "float16", "aa"),
]


@pytest.mark.parametrize("case", FAMILY_CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[4]}-{c[5]}")
def test_batch_families_and_storage_equal_single_runs(case):
    """Every collision family and storage type through the batched sweep: each member (own size and rate)
    equals its single-domain run bit for bit after 8 steps."""
    import torch
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target
    from lbmpy_amd import Timestep
    from lbmpy_amd.batch import BatchedSweep
    from lbmpy_amd.sweep import create_pdf_array
    stencil, meth, extra, faces, dtype, pattern = case
    print(f"batch family {case[:2]} {dtype} {pattern}", flush=True)
    rng = np.random.default_rng(11)
    st = LBStencil(Stencil[stencil])
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype, ghost_layers=0)
    members = []
    for _ in range(6):
        shape = tuple(int(v) for v in rng.choice((8, 16, 24, 40), 3))
        kw = dict(stencil=st, method=Method[meth], relaxation_rate=float(rng.uniform(1.5, 1.95)),
                  streaming_pattern=pattern, **extra)
        if pattern == "aa":
            ks = [create_lb_function(lbm_config=LBMConfig(**kw, timestep=t), config=cfg)
                  for t in (Timestep.ODD, Timestep.EVEN)]
        else:
            ks = [create_lb_function(lbm_config=LBMConfig(**kw), config=cfg)]
        for k in ks:
            k.set_faces(faces)
        members.append((shape, ks))
    g = torch.Generator(device="cuda").manual_seed(12)
    init = []
    for shape, _ in members:
        a = create_pdf_array(shape, st.Q, ghost_layers=0, dtype=dtype)
        a.copy_((1e-3 * torch.randn(a.shape, generator=g, device="cuda", dtype=torch.float32)).to(a.dtype))
        init.append(a)
    steps = 8
    ref = _single(members, init, steps, pattern)
    src = [a.clone() for a in init]
    dst = [torch.empty_like(a) for a in init] if pattern == "pull" else None
    b = BatchedSweep([ks[0] for _, ks in members], src, dst)
    b.run(steps)
    torch.cuda.synchronize()
    for k, (got, want) in enumerate(zip(b.current_arrays(), ref)):
        assert torch.equal(got, want), f"member {k} differs"
    assert any(not torch.equal(r, i) for r, i in zip(ref, init))
    b.close()
