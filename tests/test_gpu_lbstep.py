"""LatticeBoltzmannStep on the GPU: configurations run end to end through the step driver,
checked against the FP64 oracle (same time-loop order as lbmpy's lbstep.py:271-282) and
against the analytic Poiseuille profile (tests/poiseuille.py:119-132, rtol 0.006)."""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, ForceModel, CreateKernelConfig, Target
from lbmpy_amd.boundaries import NoSlip, slice_from_direction
from oracle.lbm_oracle import OracleDomain, equilibrium_pdfs

pytestmark = pytest.mark.gpu


def _step(domain, periodicity, lbm_config, data_type="float64", walls=("N", "S", "T", "B")):
    from lbmpy_amd.lbstep import LatticeBoltzmannStep
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type)
    step = LatticeBoltzmannStep(domain_size=domain, periodicity=periodicity, lbm_config=lbm_config, config=cfg)
    for w in walls:
        step.boundary_handling.set_boundary(NoSlip(), slice_from_direction(w, len(domain)))
    return step


def _interior_pdfs(step):
    arr = step.pdf_array
    D = step.dim
    sv = arr.permute(*range(D, -1, -1))
    return sv[(slice(None),) + (slice(1, -1),) * D].double().cpu().numpy()


def _natural(step):
    """Post-collision populations of the last step (slot i = f*_i(x))."""
    f = _interior_pdfs(step)
    if step._pattern == "aa":
        inv = step.method.stencil.inverse_indices()
        if step._aa_next_odd:        # last kernel EVEN
            return f[inv]
        raise AssertionError("compare after an even number of AA steps")
    return f


def test_config1_channel_d3q19_trt_fp64():
    """D3Q19 TRT force-driven channel, FP64, pull, 64x32x32, periodic x, NoSlip y/z,
    force 1e-7, rates [1.97, 1.6], rho=1, u=0, 500 steps."""
    shape = (64, 32, 32)
    lc = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rates=[1.97, 1.6],
                   compressible=True, force=(1e-7, 0, 0))
    step = _step(shape, (True, False, False), lc)
    assert step.boundary_handling.faces == ["periodic", "noslip", "noslip"]
    steps = 500
    step.run(steps)
    m = step.method
    dom = OracleDomain(m, shape, faces={0: "periodic", 1: "noslip", 2: "noslip"})
    rho = np.ones(tuple(reversed(shape)))
    dom.set_pdfs_interior(equilibrium_pdfs(m, rho, np.zeros((3,) + rho.shape)))
    for _ in range(steps):
        dom.step()
    err = np.abs(_natural(step) - dom.pdfs_in_natural_layout()).max()
    print(f"config-1 channel 500 steps max|GPU - oracle| = {err:.3e}", flush=True)
    assert err <= 1e-12
    # velocity getter: from the post-collision populations, u = j/rho + F/(2 rho)
    f = dom.pdfs_in_natural_layout()
    c = np.array(m.stencil.stencil_entries, float)
    r = 1 + f.sum(0)
    ux = np.einsum("q,qzyx->zyx", c[:, 0], f) / r + 0.5e-7 / r
    assert np.abs(step.velocity[:, :, :][..., 0] - ux.T).max() < 1e-15
    assert np.abs(step.density[:, :, :] - r.T).max() < 1e-14


@pytest.mark.parametrize("stencil,pattern", [(Stencil.D2Q9, "pull"), (Stencil.D3Q19, "pull"), (Stencil.D3Q19, "aa")])
def test_poiseuille_channel_gpu(stencil, pattern):
    """tests/poiseuille.py: TRT + Guo force channel, eta = 0.2, width 39 fluid cells, rtol 0.006."""
    eta = 0.2
    height = 39
    force = 0.2 / height ** 2
    st = LBStencil(stencil)
    omega = 1.0 / (3 * eta + 0.5)
    lc = LBMConfig(stencil=st, method=Method.TRT, relaxation_rate=omega, compressible=True,
                   force_model=ForceModel.GUO, force=(force, 0, 0)[:st.D], streaming_pattern=pattern)
    shape = (4, height) if st.D == 2 else (4, height, 4)
    per = (True, False) if st.D == 2 else (True, False, True)
    step = _step(shape, per, lc, walls=("N", "S"))
    last = None
    for _ in range(60):
        step.run(1000)
        ux = step.velocity[:, :] if st.D == 2 else step.velocity[:, :, :]
        rho = step.density[:, :] if st.D == 2 else step.density[:, :, :]
        # like lbmpy's step.velocity, the getter reads the post-collision populations
        # (lbstep.py:311-314): sum c f* / rho + F/(2 rho) = physical velocity + F / rho
        phys = ux[..., 0] - force / rho
        prof = phys.mean(axis=0) if st.D == 2 else phys.mean(axis=(0, 2))
        if last is not None and abs(prof.max() / last - 1) < 5e-6:
            break
        last = prof.max()
    y = np.arange(height) - (height - 1) / 2
    expected = force / (2 * eta) * (height ** 2 / 4 - y ** 2)
    np.testing.assert_allclose(prof, expected, rtol=0.006)
    uy = ux[..., 1]
    assert np.abs(uy).max() < 1e-9


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_obstacle_index_list_fp64(pattern):
    """An interior NoSlip obstacle goes through the index-list kernel (lbm355_noslip_links),
    faces through the fused walls; compared with the oracle's flag-field boundary handling."""
    shape = (16, 12, 10)
    lc = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rates=[1.8, 1.3],
                   compressible=True, force=(2e-5, 0, 0), streaming_pattern=pattern)
    step = _step(shape, (True, False, False), lc)
    obst = np.zeros(shape, dtype=bool)
    obst[6:9, 4:7, 3:6] = True
    step.boundary_handling.set_boundary(NoSlip(), mask_callback=lambda x, y, z: (x >= 6) & (x < 9) & (y >= 4) &
                                        (y < 7) & (z >= 3) & (z < 6) & (x >= 0) & (y >= 0) & (z >= 0) &
                                        (x < 16) & (y < 12) & (z < 10))
    assert len(step.boundary_handling.links) > 0
    steps = 20
    step.run(steps)
    m = step.method
    dom = OracleDomain(m, shape, faces={0: "periodic", 1: "noslip", 2: "noslip"}, pattern=pattern, obstacle=obst)
    rho = np.ones(tuple(reversed(shape)))
    f0 = equilibrium_pdfs(m, rho, np.zeros((3,) + rho.shape))
    dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
    for _ in range(steps):
        dom.step()
    fluid = ~obst.T                                   # compare fluid cells ([z, y, x] order)
    got = _natural(step)[:, fluid]
    ref = dom.pdfs_in_natural_layout()[:, fluid]
    assert np.abs(got - ref).max() <= 1e-12


@pytest.mark.parametrize("meth,pattern,dt", [(Method.TRT, "pull", "float64"), (Method.CUMULANT, "aa", "float64"),
                                             (Method.SRT, "aa", "float32"), (Method.CUMULANT, "pull", "float16")])
def test_setter_getter_roundtrip(meth, pattern, dt):
    st = LBStencil(Stencil.D3Q27)
    kw = dict(compressible=True) if meth == Method.CUMULANT else {}
    lc = LBMConfig(stencil=st, method=meth, relaxation_rate=1.7, force=(1e-5, 2e-5, 0), streaming_pattern=pattern,
                   **kw)
    shape = (10, 8, 6)
    step = _step(shape, (True, True, True), lc, data_type=dt, walls=())
    rng = np.random.default_rng(0)
    step._rho_host[...] = 1 + 0.01 * rng.standard_normal(shape)
    step._u_host[...] = 0.03 * rng.standard_normal(shape + (3,))
    rho0, u0 = step._rho_host.copy(), step._u_host.copy()
    step.set_pdf_fields_from_macroscopic_values()
    step.post_run()
    tol = {"float64": 1e-14, "float32": 1e-6, "float16": 2e-3}[dt]
    assert np.abs(step._rho_host - rho0).max() < tol
    assert np.abs(step._u_host - u0).max() < tol
    if dt == "float64":
        ref = equilibrium_pdfs(step.method, rho0.T, np.moveaxis(u0, -1, 0).transpose(0, 3, 2, 1))
        assert np.abs(_natural(step) - ref).max() < 1e-15
