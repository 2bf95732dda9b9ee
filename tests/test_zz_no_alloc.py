"""lbm355_batch_sweep_static, the allocation-free batched step, against lbm355_batch_create + lbm355_batch_sweep:
two calls in a row on one GPU (the second with a smaller ensemble, so its member table and tile map overwrite part
of what the first left in the shared scratch), each continued for several steps, both pull (step kinds 0/1) and
AA (ODD/EVEN kernels), FP32 and FP64; every member must equal the allocating path's result bit for bit.  The
entry point's limits are refused on the host without a launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _members(n, seed, pattern, dtype):
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target
    from lbmpy_amd import Timestep
    rng = np.random.default_rng(seed)
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype, ghost_layers=0)
    out = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.choice([8, 16, 24, 40], 3))
        kw = dict(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rate=float(rng.uniform(1.5, 1.9)),
                  force=(float(rng.uniform(1e-6, 1e-5)), 0.0, 0.0), streaming_pattern=pattern)
        if pattern == "aa":
            kw["timestep"] = Timestep.ODD
        k = create_lb_function(lbm_config=LBMConfig(**kw), config=cfg)
        k.set_faces(["periodic", "noslip", "periodic"])
        out.append((shape, k))
    return out


def _arrays(members, dtype, seed):
    import torch
    from lbmpy_amd.sweep import create_pdf_array
    g = torch.Generator(device="cuda").manual_seed(seed)
    res = []
    for shape, _ in members:
        a = create_pdf_array(shape, 19, ghost_layers=0, dtype=dtype)
        a.copy_((1e-3 * torch.randn(a.shape, generator=g, device="cuda", dtype=torch.float64)).to(a.dtype))
        res.append(a)
    return res


def _tables(members, src, dst):
    from lbmpy_amd import _native as nat
    n = len(members)
    descs = (nat.SweepDesc * n)()
    sp = (ctypes.c_void_p * n)()
    dp = (ctypes.c_void_p * n)()
    for i, ((_, k), s, d) in enumerate(zip(members, src, dst)):
        descs[i] = k.descriptor(s, {})
        sp[i] = s.data_ptr()
        dp[i] = d.data_ptr()
    return descs, sp, dp


def _step_kinds(pattern, steps):
    from lbmpy_amd import _native as nat
    if pattern == "pull":
        return [t % 2 for t in range(steps)]
    return [nat.AA_ODD if t % 2 == 0 else nat.AA_EVEN for t in range(steps)]


def _run_allocating(members, init, pattern, steps):
    import torch
    from lbmpy_amd import _native as nat
    lib = nat.load()
    src = [a.clone() for a in init]
    dst = [torch.empty_like(a) for a in init] if pattern == "pull" else src
    descs, sp, dp = _tables(members, src, dst)
    h = ctypes.c_void_p()
    nat.check(lib.lbm355_batch_create(descs, sp, dp, len(members), ctypes.byref(h)))
    for kind in _step_kinds(pattern, steps):
        nat.check(lib.lbm355_batch_sweep(h, kind, 256, 1 << 20, 0, nat.current_stream_handle()))
    torch.cuda.synchronize()
    nat.check(lib.lbm355_batch_destroy(h))
    return src if steps % 2 == 0 or pattern != "pull" else dst


def _run_static(members, init, pattern, steps):
    import torch
    from lbmpy_amd import _native as nat
    lib = nat.load()
    src = [a.clone() for a in init]
    dst = [torch.empty_like(a) for a in init] if pattern == "pull" else src
    descs, sp, dp = _tables(members, src, dst)
    for kind in _step_kinds(pattern, steps):
        nat.check(lib.lbm355_batch_sweep_static(descs, sp, dp, len(members), kind, 256, 1 << 20, 0,
                                                nat.current_stream_handle()))
    torch.cuda.synchronize()
    return src if steps % 2 == 0 or pattern != "pull" else dst


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_static_batch_equals_allocating_batch_on_two_calls(pattern, dtype):
    import torch
    print(f"static batch {pattern} {dtype}", flush=True)
    for n, seed in ((9, 1), (4, 2)):                 # the second ensemble is smaller
        members = _members(n, seed, pattern, dtype)
        init = _arrays(members, dtype, seed + 10)
        steps = 6
        want = _run_allocating(members, init, pattern, steps)
        got = _run_static(members, init, pattern, steps)
        for i, (g, w) in enumerate(zip(got, want)):
            assert torch.equal(g, w), f"member {i} of {n} differs"
        assert any(float((g - i0).abs().max()) > 0 for g, i0 in zip(got, init))


def test_static_batch_limits_are_refused():
    from lbmpy_amd import _native as nat
    lib = nat.load()
    members = _members(2, 3, "pull", "float32")
    import torch
    src = _arrays(members, "float32", 4)
    dst = [torch.empty_like(a) for a in src]
    descs, sp, dp = _tables(members, src, dst)
    assert lib.lbm355_batch_sweep_static(descs, sp, dp, 2, 0, 100, 4, 0, nat.current_stream_handle()) != 0
    assert "threads" in lib.lbm355_last_error().decode()
    assert lib.lbm355_batch_sweep_static(descs, sp, dp, 2, 5, 256, 4, 0, nat.current_stream_handle()) != 0
    big = (nat.SweepDesc * 1025)()
    bsp = (ctypes.c_void_p * 1025)()
    bdp = (ctypes.c_void_p * 1025)()
    for i in range(1025):
        big[i] = descs[0]
        bsp[i] = sp[0]
        bdp[i] = dp[0]
    assert lib.lbm355_batch_sweep_static(big, bsp, bdp, 1025, 0, 256, 4, 0, nat.current_stream_handle()) != 0
    assert "1024" in lib.lbm355_last_error().decode()
