"""Occupancy caps: a launch's waves_per_simd (0 the library's table, -1 none, 1..8 a cap) only changes how many
workgroups share a CU, never a result; tune_occupancy times the caps on a copy of the state and keeps the fastest."""
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, CreateKernelConfig, Target, Timestep, create_lb_function
from lbmpy_amd.sweep import create_pdf_array, tune_occupancy

pytestmark = pytest.mark.gpu

CAPS = (0, -1, 1, 2, 3, 5, 8)


def _kernels(stencil, pattern, dtype="float32", cfg_kw=None, **kw):
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype, ghost_layers=0, **(cfg_kw or {}))
    kw = kw or dict(method=Method.SRT, relaxation_rate=1.8)
    base = dict(stencil=LBStencil(stencil), streaming_pattern=pattern, **kw)
    if pattern in ("pull", "push"):
        ks = [create_lb_function(lbm_config=LBMConfig(**base), config=cfg)]
    else:
        ks = [create_lb_function(lbm_config=LBMConfig(**base, timestep=t), config=cfg) for t in (Timestep.ODD, Timestep.EVEN)]
    for k in ks:
        k.set_faces(["periodic"] * 3)
    return ks


def _random_state(shape, q, dtype, seed=3):
    g = torch.Generator(device="cuda").manual_seed(seed)
    a = create_pdf_array(shape, q, ghost_layers=0, dtype=dtype)
    a.copy_((torch.randn(a.shape, generator=g, device="cuda", dtype=torch.float32) * 1e-3).to(a.dtype))
    return a


def _run(ks, a0, steps):
    a = a0.clone()
    b = a0.clone() if len(ks) == 1 else None
    for t in range(steps):
        if b is not None:
            ks[0](src=a if t % 2 == 0 else b, dst=b if t % 2 == 0 else a)
        else:
            ks[t % 2](src=a)
    torch.cuda.synchronize()
    return b if (b is not None and steps % 2) else a


@pytest.mark.parametrize("stencil,pattern,kw", [
    (Stencil.D3Q27, "aa", dict(method=Method.CUMULANT, relaxation_rate=1.99, compressible=True)),
    (Stencil.D3Q19, "pull", {}),
    (Stencil.D3Q19, "esotwist", {}),
    (Stencil.D3Q27, "aa", dict(method=Method.CENTRAL_MOMENT, relaxation_rate=1.8, compressible=True)),
])
def test_caps_do_not_change_results(stencil, pattern, kw):
    ks = _kernels(stencil, pattern, **kw)
    q = LBStencil(stencil).Q
    a0 = _random_state((128, 24, 20), q, "float32")
    ref = None
    for cap in CAPS:
        for k in ks:
            k.waves_per_simd = cap
        print(f"{stencil} {pattern} waves_per_simd={cap}", flush=True)
        out = _run(ks, a0, 5)
        if ref is None:
            ref = out
        else:
            assert torch.equal(out, ref), f"waves_per_simd={cap}"


def test_bad_cap_is_refused():
    k = _kernels(Stencil.D3Q19, "pull")[0]
    a = create_pdf_array((64, 8, 8), 19, ghost_layers=0, dtype="float32")
    b = torch.empty_like(a)
    for cap in (9, -2):
        k.waves_per_simd = cap
        with pytest.raises(Exception):
            k(src=a, dst=b)


def test_tune_occupancy_keeps_the_state_and_picks_a_candidate():
    ks = _kernels(Stencil.D3Q27, "aa", method=Method.CUMULANT, relaxation_rate=1.99, compressible=True)
    a = _random_state((128, 64, 64), 27, "float32")
    before = a.clone()
    cands = (0, -1, 2, 3)
    t = tune_occupancy(ks, a, candidates=cands, reps=4)
    torch.cuda.synchronize()
    assert torch.equal(a, before)
    assert set(t) == {0, 1}
    for i, k in enumerate(ks):
        assert set(t[i]) == set(cands)
        assert k.waves_per_simd == min(t[i], key=t[i].get)
    print("tuned caps", [k.waves_per_simd for k in ks], t, flush=True)


def test_tune_occupancy_keeps_the_rounding_sequence():
    sr = dict(stochastic_rounding=True, rounding_seed=11)
    ks = _kernels(Stencil.D3Q19, "pull", dtype="float16", cfg_kw=sr)
    ref = _kernels(Stencil.D3Q19, "pull", dtype="float16", cfg_kw=sr)
    a0 = _random_state((64, 16, 16), 19, "float16")
    tune_occupancy(ks, a0, a0.clone(), candidates=(0, 2), reps=2)
    assert ks[0].rounding_launches == 0
    assert torch.equal(_run(ks, a0, 3), _run(ref, a0, 3))
