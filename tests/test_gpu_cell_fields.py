"""Per-cell fields beside the PDFs in the GPU sweep, against the FP64 oracle:

* the Smagorinsky model's ``omega_output_field`` and ``eddy_viscosity_field`` (turbulence_models.py:73-81): the
  sweep writes each cell's shear rate omega_sgs and Cs^2 * 3/2 * omega_sgs * Pi, computed from the populations
  it collides; pull and AA (both kernels), to 1e-13;
* ``density_input`` / ``velocity_input`` (creationfunctions.py:655-680) for the SRT / TRT / MRT sweep: the
  equilibrium and the force term take the given fields, conserved moments relax (conserved_moments = False);
  pull and AA over 6 steps, to 1e-12;
* the fused ``output`` fields (creationfunctions.py:299-305) equal the macroscopic getter of the populations the
  sweep leaves (no force: collision keeps density and momentum), pull and AA;
* field tensors whose shape or layout the kernel cannot address are refused on the host.
"""
import numpy as np
import pytest
import torch

from lbmpy_amd import (LBMConfig, LBStencil, Stencil, Method, ForceModel, SubgridScaleModel, create_lb_method,
                       Timestep)
from lbmpy_amd.macroscopic_value_kernels import macroscopic_values_getter
from lbmpy_amd.moments import is_shear_moment
from lbmpy_amd.sweep import create_pdf_array
from oracle.lbm_oracle import OracleDomain
from tests.lbm_helpers import seeded_state, storage_view, gpu_kernels, gpu_run

pytestmark = pytest.mark.gpu


def _cell_field(shape, components=0, dtype=torch.float64, g=1):
    """A zeroed [x, y, (z,) (c)] tensor over the PDF array's spatial extent, x contiguous (fzyx)."""
    full = tuple(n + 2 * g for n in shape)
    if components:
        return torch.zeros((components,) + full[::-1], dtype=dtype, device="cuda").permute(
            *range(len(shape), 0, -1), 0)
    return torch.zeros(full[::-1], dtype=dtype, device="cuda").permute(*range(len(shape) - 1, -1, -1))


def _interior_cells(t, D):
    """Interior values of a scalar [x, y, (z)] field in the oracle's cell order (z, y, x flattened)."""
    inner = t[(slice(1, -1),) * D].cpu().numpy()
    return inner.T.reshape(-1)


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_smagorinsky_rate_and_eddy_viscosity_fields(pattern):
    cs = 0.12
    shape = (8, 6, 5)
    D = 3
    omega_f, nut_f = _cell_field(shape), _cell_field(shape)
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.9, compressible=True,
              streaming_pattern=pattern, subgrid_scale_model=(SubgridScaleModel.SMAGORINSKY, cs),
              force=(1e-5, 0, 0), omega_output_field=omega_f, eddy_viscosity_field=nut_f)
    m = create_lb_method(LBMConfig(**{k: v for k, v in kw.items()
                                      if k not in ("omega_output_field", "eddy_viscosity_field")}))
    f0 = seeded_state(m, shape, seed=21, du=0.05, noise=3e-3)
    faces = {a: "periodic" for a in range(D)}
    # the oracle's per-cell rate and eddy viscosity from the populations each pull step collides
    dom = OracleDomain(m, shape, faces=faces, pattern="pull", collision={"smagorinsky": cs})
    dom.set_pdfs_interior(f0)
    shear = next(i for i, p in enumerate(dom.coll.polys) if is_shear_moment(p, D))
    want = []
    for _ in range(2):
        fpre = dom.pull_pre_collision().reshape(27, -1)
        rho, drho, u, _ = dom.coll.macroscopic(fpre)
        rates = np.repeat(dom.coll.rates[:, None], fpre.shape[1], axis=1)
        w = dom.coll._smagorinsky_rates(fpre, rho, drho, u, rates)[shear]
        p = dom.coll._pi_neq(fpre, rho, drho, u)
        pi = np.sqrt(2.0 * ((p[:3] ** 2).sum(0) + 2.0 * (p[3:] ** 2).sum(0))) / rho
        want.append((w, 1.5 * cs ** 2 * w * pi))
        dom.step()
    assert np.abs(want[0][0] - 1.9).max() > 1e-3                      # the subgrid model is active
    kernels = gpu_kernels(kw)
    print(f"smagorinsky fields {pattern}", flush=True)
    for steps in (1, 2):
        omega_f.zero_()
        nut_f.zero_()
        if pattern == "aa" and steps == 1:
            # one ODD step from the AA layout (gpu_run compares AA runs after EVEN kernels only)
            even, odd = kernels
            src = create_pdf_array(shape, 27, 1, "float64")
            sl = (slice(None),) + (slice(1, -1),) * D
            storage_view(src, D)[sl] = torch.from_numpy(f0[np.array(m.stencil.inverse_indices())]).cuda()
            for k in kernels:
                k.set_faces(["periodic"] * D)
            odd(src=src)
        else:
            gpu_run(kernels, m, shape, faces, f0, steps)
        torch.cuda.synchronize()
        w_want, nut_want = want[steps - 1]
        assert np.abs(_interior_cells(omega_f, D) - w_want).max() <= 1e-13
        assert np.abs(_interior_cells(nut_f, D) - nut_want).max() <= 1e-16 + 1e-12 * np.abs(nut_want).max()


def test_subgrid_outputs_need_the_smagorinsky_model():
    from lbmpy_amd import create_lb_function, CreateKernelConfig, Target
    with pytest.raises(NotImplementedError):
        create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT,
                                                relaxation_rate=1.9, compressible=True,
                                                omega_output_field=_cell_field((4, 4, 4))),
                           config=CreateKernelConfig(target=Target.GPU))


INPUT_CASES = [
    ("D2Q9", dict(method=Method.SRT, relaxation_rate=1.7, compressible=True), (12, 9)),
    ("D2Q9", dict(method=Method.TRT, relaxation_rates=[1.6, 1.2], compressible=False, force=(1e-4, 2e-5),
                  force_model=ForceModel.LUO), (10, 7)),
    ("D3Q19", dict(method=Method.MRT, relaxation_rate=1.8, compressible=True, force=(1e-5, 0, 2e-5),
                   force_model=ForceModel.GUO), (6, 5, 4)),
    ("D3Q19", dict(method=Method.TRT, relaxation_rate=1.75, compressible=False), (5, 6, 4)),
]


@pytest.mark.parametrize("with_density", [True, False], ids=["rho+u", "u"])
@pytest.mark.parametrize("pattern", ["pull", "aa"])
@pytest.mark.parametrize("case", INPUT_CASES, ids=lambda c: f"{c[0]}-{c[1]['method'].name}")
def test_density_and_velocity_inputs_match_oracle(case, pattern, with_density):
    st, extra, shape = case
    D = len(shape)
    rng = np.random.default_rng(31)
    grid = tuple(reversed(shape))
    rho_in = 1.0 + 0.02 * rng.standard_normal(grid)
    u_in = 0.04 * rng.standard_normal((D,) + grid)
    rho_t = _cell_field(shape)
    u_t = _cell_field(shape, D)
    inner = (slice(1, -1),) * D
    rho_t[inner] = torch.from_numpy(np.ascontiguousarray(rho_in.T)).cuda()
    for a in range(D):
        u_t[inner + (a,)] = torch.from_numpy(np.ascontiguousarray(u_in[a].T)).cuda()
    kw = dict(stencil=LBStencil(Stencil[st]), streaming_pattern=pattern, velocity_input=u_t, **extra)
    if with_density:
        kw["density_input"] = rho_t
    m = create_lb_method(LBMConfig(**{k: v for k, v in kw.items() if k not in ("velocity_input", "density_input")},
                                   conserved_moments=False))
    f0 = seeded_state(m, shape, seed=32, du=0.03, noise=1e-3)
    faces = {a: "periodic" for a in range(D)}
    inputs = (rho_in.reshape(-1) if with_density else None, u_in.reshape(D, -1))
    dom = OracleDomain(m, shape, faces=faces, pattern=pattern, collision={"inputs": inputs})
    dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
    steps = 6
    for _ in range(steps):
        dom.step()
    ref = dom.pdfs_in_natural_layout()
    print(f"inputs {st} {extra['method'].name} {pattern} density={with_density}", flush=True)
    got = gpu_run(gpu_kernels(kw), m, shape, faces, f0, steps)
    assert np.abs(got - ref).max() <= 1e-12
    # the inputs drive the flow away from what the populations alone would give
    plain = OracleDomain(m, shape, faces=faces, pattern=pattern)
    plain.set_pdfs_interior(f0[plain.inv] if pattern == "aa" else f0)
    for _ in range(steps):
        plain.step()
    assert np.abs(plain.pdfs_in_natural_layout() - ref).max() > 1e-4


def test_inputs_are_refused_for_cumulants_and_without_velocity():
    from lbmpy_amd import create_lb_function, CreateKernelConfig, Target
    cfg = CreateKernelConfig(target=Target.GPU)
    u_t = _cell_field((4, 4, 4), 3)
    with pytest.raises(NotImplementedError):
        create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT,
                                                relaxation_rate=1.9, compressible=True, velocity_input=u_t),
                           config=cfg)
    with pytest.raises(ValueError):
        create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.SRT,
                                                relaxation_rate=1.9, density_input=_cell_field((4, 4, 4))),
                           config=cfg)


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_fused_outputs_equal_the_getter(pattern):
    shape = (9, 7, 6)
    D = 3
    rho_o, u_o = _cell_field(shape), _cell_field(shape, D)
    kw = dict(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rates=[1.7, 1.3], compressible=True,
              streaming_pattern=pattern, output={"density": rho_o, "velocity": u_o})
    m = create_lb_method(LBMConfig(**{k: v for k, v in kw.items() if k != "output"}))
    f0 = seeded_state(m, shape, seed=41, du=0.04, noise=1e-3)
    kernels = gpu_kernels(kw)
    for k in kernels:
        k.set_faces(["periodic"] * D)
    src = create_pdf_array(shape, 19, 1, "float64")
    sl = (slice(None),) + (slice(1, -1),) * D
    first = f0[np.array(m.stencil.inverse_indices())] if pattern == "aa" else f0
    storage_view(src, D)[sl] = torch.from_numpy(first).cuda()
    dst = create_pdf_array(shape, 19, 1, "float64") if pattern == "pull" else src
    for t in range(3):
        if pattern == "pull":
            kernels[0](src=src, dst=dst)
            src, dst = dst, src
            prev = Timestep.BOTH
        else:
            even, odd = kernels
            (odd if t % 2 == 0 else even)(src=src)
            prev = Timestep.ODD if t % 2 == 0 else Timestep.EVEN
        r2, v2 = _cell_field(shape), _cell_field(shape, D)
        macroscopic_values_getter(m, r2, v2, src, streaming_pattern=pattern, previous_timestep=prev,
                                  periodicity=(True,) * D).compile()()
        inner = (slice(1, -1),) * D
        assert torch.abs(rho_o[inner] - r2[inner]).max().item() <= 1e-14
        assert torch.abs(u_o[inner] - v2[inner]).max().item() <= 1e-14
        assert torch.abs(u_o[inner]).max().item() > 1e-3


def test_cell_fields_the_kernel_cannot_address_are_refused():
    from lbmpy_amd import create_lb_function, CreateKernelConfig, Target
    shape = (6, 5, 4)
    src = create_pdf_array(shape, 19, 1, "float64")
    bad = {
        "small": {"density": torch.zeros((6, 5, 4), dtype=torch.float64, device="cuda")},
        "few_components": {"velocity": _cell_field(shape, 2)},
        "fp32": {"density": _cell_field(shape, dtype=torch.float32)},
        "x_strided": {"density": torch.zeros((8, 7, 6), dtype=torch.float64, device="cuda")},
    }
    for name, out in bad.items():
        k = create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.SRT,
                                                    relaxation_rate=1.8, output=out),
                               config=CreateKernelConfig(target=Target.GPU))
        k.set_faces(["periodic"] * 3)
        with pytest.raises(ValueError):
            k(src=src, dst=create_pdf_array(shape, 19, 1, "float64"))
    with pytest.raises(NotImplementedError):
        create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.SRT,
                                                relaxation_rate=1.8, output={"moment2": _cell_field(shape)}),
                           config=CreateKernelConfig(target=Target.GPU))
