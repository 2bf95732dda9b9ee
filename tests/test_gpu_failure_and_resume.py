"""Failure detection and checkpoint / resume on the GPU.

* lbm355_count_nonfinite counts NaN and infinite populations; lbmpy's quick test of an unstable SRT cavity
  (tests/test_quicktests.py:15-27: D2Q9, omega 1.9999, lid 0.05, 1000 steps) must be caught, and its cumulant
  cavity (:30-35) must stay finite.
* A run saved and resumed continues bit for bit: AA streaming saved between an odd and an even step, FP16 storage
  with stochastic rounding (whose random stream follows the launch count), and pull streaming."""
import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, ForceModel, CreateKernelConfig, Target

pytestmark = pytest.mark.gpu


def test_nonfinite_count_sees_injected_values():
    from lbmpy_amd.scenarios import create_fully_periodic_flow
    print("non-finite count: injected NaN and Inf", flush=True)
    n = 12
    u = np.zeros((n, n, n, 3))
    u[..., 0] = 0.02 * np.sin(2 * np.pi * np.arange(n) / n)[None, :, None]
    step = create_fully_periodic_flow(u, lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT,
                                                              relaxation_rate=1.7))
    step.run(4)
    assert step.nonfinite_count() == 0
    pdf = step.pdf_array
    pdf[3, 4, 5, 7] = float("nan")
    pdf[1, 2, 3, 0] = float("inf")
    assert step.nonfinite_count() == 2
    with pytest.raises(FloatingPointError):
        step.check_finite()


def test_unstable_srt_cavity_is_caught_and_cumulant_cavity_stays_finite():
    from lbmpy_amd.scenarios import create_lid_driven_cavity
    print("non-finite count: lbmpy's unstable SRT cavity", flush=True)
    srt = create_lid_driven_cavity((40, 40), relaxation_rate=1.9999, lid_velocity=0.05, compressible=True,
                                   zero_centered=False, force=(-1e-10, 0), force_model=ForceModel.LUO)
    srt.run(1000)
    assert srt.nonfinite_count() > 0
    assert np.isnan(np.max(srt.velocity[:, :]))
    cum = create_lid_driven_cavity((40, 40, 40), method=Method.CUMULANT, relaxation_rate=1.999999, compressible=True,
                                   force=(-1e-10, 0, 0))
    cum.run(100)
    assert cum.nonfinite_count() == 0
    cum.check_finite()


def _bits(t):
    return t.contiguous().view(torch.uint8)


def _channel(pattern, data_type, sr):
    from lbmpy_amd.scenarios import create_channel
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type, compute_type="float32",
                             stochastic_rounding=sr, rounding_seed=11)
    return create_channel((48, 24, 8), force=2e-5, duct=True, stencil=LBStencil(Stencil.D3Q19), method=Method.MRT,
                          relaxation_rate=1.8, streaming_pattern=pattern, config=cfg)


@pytest.mark.parametrize("pattern,data_type,sr", [("aa", "float32", False), ("aa", "float16", True),
                                                  ("pull", "float16", True), ("pull", "float64", False)])
def test_resume_continues_bit_for_bit(tmp_path, pattern, data_type, sr):
    print(f"checkpoint/resume: {pattern} {data_type} stochastic_rounding={sr}", flush=True)
    a = _channel(pattern, data_type, sr)
    a.run(31)                                   # odd: an AA field between an odd and an even step
    a.save_checkpoint(tmp_path / "run.pt")
    a.run(20)
    want = a.pdf_array.clone()
    b = _channel(pattern, data_type, sr)
    b.load_checkpoint(tmp_path / "run.pt")
    assert b.time_steps_run == 31
    b.run(20)
    got = b.pdf_array
    assert torch.equal(_bits(got), _bits(want))
    # without the saved state the same continuation differs (parity or random stream)
    c = _channel(pattern, data_type, sr)
    c.pdf_array.copy_(torch.load(tmp_path / "run.pt", weights_only=True)["pdf"].cuda())
    c.run(20)
    if pattern == "aa" or sr:
        assert not torch.equal(_bits(c.pdf_array), _bits(want))
