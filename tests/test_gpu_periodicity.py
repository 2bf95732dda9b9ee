"""Ghost-layer periodicity on the GPU (lbm355_copy_boxes behind LBMPeriodicityHandling,
advanced_streaming/communication.py:21-126):

* the copies equal lbmpy's sequential slice assignments, bit for bit, for every stencil, pattern, parity and
  storage type;
* sweeps with ghost faces plus the handler, stepped in the order of lbmpy's tests
  (tests/advanced_streaming/test_fully_periodic_flow.py:104-119), feed the next kernel bit-identical values
  (every interior cell and direction, read without wrap) to the same sweeps with periodicity built into their
  index math (addresses wrapped), after every step (FP64 and FP32);
* lbmpy's fully periodic flow test: a uniform flow stays uniform for 100 steps with every pattern, and every
  pattern gives the same velocity field.
"""
import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, create_lb_method, \
    CreateKernelConfig, Target
from lbmpy_amd.advanced_streaming import (LBMPeriodicityHandling, GpuArrays, get_timesteps, is_inplace,
                                          streaming_patterns)
from lbmpy_amd.advanced_streaming.utility import access_table, kernel_code
from lbmpy_amd.macroscopic_value_kernels import macroscopic_values_getter, macroscopic_values_setter
from lbmpy_amd.sweep import create_pdf_array

pytestmark = pytest.mark.gpu


def _random_pdfs(shape, Q, dtype, seed):
    arr = create_pdf_array(shape, Q, 1, dtype)
    g = torch.Generator().manual_seed(seed)
    arr.copy_(torch.rand(tuple(arr.shape), generator=g, dtype=torch.float64).to(arr.dtype))
    return arr


@pytest.mark.parametrize("dtype", ["float64", "float32", "float16"])
@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
def test_copies_equal_slice_assignments(stencil, pattern, dtype):
    st = LBStencil(Stencil[stencil])
    shape = (5, 4) if st.D == 2 else (5, 4, 3)
    arr = _random_pdfs(shape, st.Q, dtype, seed=5)
    h = LBMPeriodicityHandling(st, GpuArrays((True,) * st.D, pdfs=arr), "pdfs", streaming_pattern=pattern)
    for t in get_timesteps(pattern):
        ref = arr.cpu()
        for src, dst in h.comm_slices[t.idx]:
            ref[dst] = ref[src]
        h(t)
        torch.cuda.synchronize()
        assert torch.equal(arr.cpu(), ref), (stencil, pattern, t)


def test_copies_large_domain_many_blocks():
    st = LBStencil(Stencil.D3Q27)
    arr = _random_pdfs((64, 32, 40), 27, "float32", seed=6)
    h = LBMPeriodicityHandling(st, GpuArrays((True, True, True), pdfs=arr), "pdfs", streaming_pattern="esopush")
    for t in get_timesteps("esopush"):
        ref = arr.cpu()
        for src, dst in h.comm_slices[t.idx]:
            ref[dst] = ref[src]
        h(t)
        torch.cuda.synchronize()
        assert torch.equal(arr.cpu(), ref)


CASES = [
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6, force=(1e-4, 2e-5)), (9, 6)),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.8, 1.2]), (7, 5, 6)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.9, compressible=True, force=(1e-5, 0, 0)), (6, 5, 4)),
]


def _kernels(kw, pattern, dtype, periodic):
    cfg = CreateKernelConfig(target=Target.GPU, data_type=dtype, ghost_layers=1)
    if is_inplace(pattern):
        ks = [create_lb_function(lbm_config=LBMConfig(**kw, streaming_pattern=pattern, timestep=t), config=cfg)
              for t in get_timesteps(pattern)]
    else:
        ks = [create_lb_function(lbm_config=LBMConfig(**kw, streaming_pattern=pattern), config=cfg)]
    D = kw["stencil"].D
    for k in ks:
        k.set_faces(["periodic" if periodic else "ghost"] * D)
    return ks


def _unwrap(arr, st, shape, code):
    """The same state with the previous kernel's writes not wrapped: each value sits where a kernel without
    built-in periodicity would have written it (in the ghost layer where the write leaves the domain)."""
    out = arr.clone()
    tab = access_table(st, code)
    D = st.D
    cells = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), axis=-1).reshape(-1, D)
    for i in range(st.Q):
        tgt = cells + tab[i, 4:4 + D]
        wrapped = tgt % np.array(shape)
        slot = int(tab[i, 7])
        dst = tuple(torch.from_numpy(tgt[:, a] + 1).cuda() for a in range(D)) + (slot,)
        src = tuple(torch.from_numpy(wrapped[:, a] + 1).cuda() for a in range(D)) + (slot,)
        out[dst] = arr[src]
    return out


def _next_reads(arr, st, shape, code, wrap):
    """[Q, cells]: the values the kernel `code` reads for every interior cell and direction, addresses wrapped
    into the interior (built-in periodicity) or not (ghost layers)."""
    tab = access_table(st, code)
    D = st.D
    cells = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), axis=-1).reshape(-1, D)
    out = []
    for i in range(st.Q):
        src = cells + tab[i, :D]
        if wrap:
            src = src % np.array(shape)
        out.append(arr[tuple(torch.from_numpy(src[:, a] + 1).cuda() for a in range(D)) + (int(tab[i, 3]),)])
    return torch.stack(out)


@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_ghost_faces_with_handler_equal_builtin_periodicity(case, pattern):
    stencil, meth, extra, shape = case
    st = LBStencil(Stencil[stencil])
    kw = dict(stencil=st, method=meth, **extra)
    D, Q = st.D, st.Q
    for dtype in (["float64", "float32"] if stencil == "D3Q19" else ["float64"]):
        per_k = _kernels(kw, pattern, dtype, True)
        gh_k = _kernels(kw, pattern, dtype, False)
        zeroth = get_timesteps(pattern)[0]
        a = _random_pdfs(shape, Q, dtype, seed=9)
        a.mul_(0.02).add_(1.0 / Q)
        b = _unwrap(a, st, shape, kernel_code(pattern, zeroth))
        a_tmp, b_tmp = torch.zeros_like(a), torch.zeros_like(b)
        dh = GpuArrays((True,) * D, pdfs=b, pdfs_tmp=b_tmp)
        handler = LBMPeriodicityHandling(st, dh, "pdfs", streaming_pattern=pattern)
        cur = zeroth
        for step in range(5):
            handler(cur)
            nxt = kernel_code(pattern, cur.next())
            assert torch.equal(_next_reads(a, st, shape, nxt, True),
                               _next_reads(dh.gpu_arrays["pdfs"], st, shape, nxt, False)), (pattern, dtype, step)
            cur = cur.next()
            if is_inplace(pattern):
                per_k[cur.idx](src=a)
                gh_k[cur.idx](src=dh.gpu_arrays["pdfs"])
            else:
                per_k[0](src=a, dst=a_tmp)
                a, a_tmp = a_tmp, a
                gh_k[0](src=dh.gpu_arrays["pdfs"], dst=dh.gpu_arrays["pdfs_tmp"])
                dh.swap("pdfs", "pdfs_tmp")
        handler(cur)
        nxt = kernel_code(pattern, cur.next())
        final = _next_reads(a, st, shape, nxt, True)
        assert torch.equal(final, _next_reads(dh.gpu_arrays["pdfs"], st, shape, nxt, False))
        assert (final.double() - (1.0 / Q)).abs().max() > 1e-3


_FLOWS = {}


@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
def test_fully_periodic_flow(stencil, pattern):
    st = LBStencil(Stencil[stencil])
    D = st.D
    domain = (30,) * D
    cfg = LBMConfig(stencil=st, method=Method.SRT, relaxation_rate=1.0, streaming_pattern=pattern)
    method = create_lb_method(cfg)
    ks = _kernels(dict(stencil=st, method=Method.SRT, relaxation_rate=1.0), pattern, "float64", False)
    pdfs = create_pdf_array(domain, st.Q, 1, "float64")
    dh = GpuArrays((True,) * D, pdfs=pdfs, pdfs_tmp=torch.zeros_like(pdfs))
    u_x = 0.01
    velocity = torch.zeros(tuple(pdfs.shape[:-1]) + (D,), dtype=torch.float64, device=pdfs.device)
    timesteps = get_timesteps(pattern)
    zeroth = timesteps[0]
    setter = macroscopic_values_setter(method, 1.0, (u_x,) * D, pdfs, streaming_pattern=pattern,
                                       previous_timestep=zeroth).compile()
    getters = [macroscopic_values_getter(method, None, velocity, pdfs, streaming_pattern=pattern,
                                         previous_timestep=t).compile() for t in timesteps]
    handler = LBMPeriodicityHandling(st, dh, "pdfs", streaming_pattern=pattern)
    setter()
    cur = zeroth
    for _ in range(100):
        handler(cur)
        cur = cur.next()
        if is_inplace(pattern):
            ks[cur.idx](src=dh.gpu_arrays["pdfs"])
        else:
            ks[0](src=dh.gpu_arrays["pdfs"], dst=dh.gpu_arrays["pdfs_tmp"])
            dh.swap("pdfs", "pdfs_tmp")
        getters[cur.idx](pdfs=dh.gpu_arrays["pdfs"])
    u = velocity[tuple(slice(1, -1) for _ in range(D))].cpu().numpy()
    np.testing.assert_allclose(u, np.full(domain + (D,), u_x))
    for (st_name, other), prev_u in _FLOWS.items():
        if st_name == stencil:
            np.testing.assert_allclose(u, prev_u, err_msg=f"{pattern} differs from {other}")
    _FLOWS[(stencil, pattern)] = u
