"""Taylor-Green vortex at 64^3 (D3Q27 cumulant, AA, zero-centred, fully periodic, u0 = 0.05, Re = 1600),
2000 steps on the GPU against the C oracle's run of the same setup (tests/golden/tgv64_cumulant_aa.npz).

FP64: sampled PDFs within 1e-12 max-abs after 500 and 2000 steps; per-plane population sums (each a sum
of 4096 PDFs) within 1e-10.
FP32 (FP32 storage and arithmetic): velocity at the sampled cells within 1e-5 of u0 after 2000 steps,
density within 1e-6.
FP16 storage (FP32 arithmetic): velocity within 5e-3 of u0, density within 1e-4 (FP16 keeps 11 bits of
the zero-centred deviations).
"""
import json
import os

import numpy as np
import pytest

from tests.golden.make_tgv64 import OUT, N, SAMPLE, tgv_method, tgv_initial_pdfs

pytestmark = pytest.mark.gpu


def _fixture():
    if not os.path.exists(OUT):
        pytest.skip("64^3 TGV fixture missing (python tests/golden/make_tgv64.py)")
    z = np.load(OUT, allow_pickle=False)
    return json.loads(str(z["meta"])), z


def _macroscopic(sample, method):
    c = np.array(method.stencil.stencil_entries, dtype=np.float64)
    rho = 1.0 + sample.sum(axis=0)
    u = np.einsum("qa,q...->a...", c, sample) / rho
    return rho, u


def _run(dtype, snapshots):
    import torch
    from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target,
                           Timestep)
    from lbmpy_amd.sweep import create_pdf_array
    m, omega = tgv_method()
    f0 = tgv_initial_pdfs(m)
    inv = np.array(m.stencil.inverse_indices())
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype,
                             compute_type="float64" if dtype == "float64" else "float32", ghost_layers=0)
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=omega, compressible=True,
              zero_centered=True, streaming_pattern="aa")
    even = create_lb_function(lbm_config=LBMConfig(**kw, timestep=Timestep.EVEN), config=cfg)
    odd = create_lb_function(lbm_config=LBMConfig(**kw, timestep=Timestep.ODD), config=cfg)
    for k in (even, odd):
        k.set_faces(["periodic"] * 3)
    pdf = create_pdf_array((N, N, N), 27, ghost_layers=0, dtype=dtype)
    storage = pdf.permute(3, 2, 1, 0)
    storage.copy_(torch.from_numpy(f0[inv]).to(storage.dtype).cuda())
    out = {}
    for t in range(max(snapshots)):
        (odd if t % 2 == 0 else even)(src=pdf)
        if t + 1 in snapshots:
            f = storage.double().cpu().numpy()[inv]
            out[t + 1] = (f[:, ::SAMPLE, ::SAMPLE, ::SAMPLE].copy(), f.sum(axis=(2, 3)).T.copy())
    return m, out


def test_tgv64_fp64_matches_oracle():
    meta, z = _fixture()
    snaps = tuple(meta["snapshots"])
    print("tgv64 fp64", flush=True)
    _, got = _run("float64", snaps)
    for s in snaps:
        err = np.abs(got[s][0] - z[f"sample_{s}"]).max()
        perr = np.abs(got[s][1] - z[f"planes_{s}"]).max()
        print(f"step {s}: sampled PDFs max |GPU - oracle| = {err:.3e}, plane sums {perr:.3e}", flush=True)
        assert err <= 1e-12
        assert perr <= 1e-10


@pytest.mark.parametrize("dtype,u_tol,rho_tol", [("float32", 1e-5, 1e-6), ("float16", 5e-3, 1e-4)])
def test_tgv64_reduced_precision_error(dtype, u_tol, rho_tol):
    meta, z = _fixture()
    print(f"tgv64 {dtype}", flush=True)
    m, got = _run(dtype, (2000,))
    rho_ref, u_ref = _macroscopic(z["sample_2000"], m)
    rho, u = _macroscopic(got[2000][0], m)
    eu = np.abs(u - u_ref).max() / meta["u0"]
    er = np.abs(rho - rho_ref).max()
    print(f"{dtype} after 2000 steps: max |u - u_oracle| / u0 = {eu:.3e}, max |rho - rho_oracle| = {er:.3e}",
          flush=True)
    assert np.isfinite(u).all()
    assert eu <= u_tol
    assert er <= rho_tol
