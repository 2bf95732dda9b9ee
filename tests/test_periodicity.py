"""Ghost-layer periodicity (lbmpy advanced_streaming/communication.py) without a GPU:

* the kernels' access tables (lbm355_pattern_access, what the HIP sweeps are compiled with) equal the oracle's
  restatement of lbmpy's accessors for every stencil and kernel;
* lbmpy's own slice tests (tests/advanced_streaming/test_communication.py:20-88): slices are non-empty,
  source and destination have one shape, and pull copies the layer before the ghost layer onto the
  opposite ghost region;
* the closed form the copies exist for: a kernel writing without wrap into a ghost-layer array, followed by
  the copies, leaves every value the next kernel reads without wrap equal to what it reads on a periodic
  domain (addresses wrapped), for every pattern and parity, fully and partly periodic;
* the boxes handed to lbm355_copy_boxes copy exactly what the slices copy (a CPU emulation of the kernel),
  and a call's sources and destinations are disjoint, so concurrent copies equal lbmpy's sequential ones;
* a planted mistake (one slice pair dropped) fails the closed form.
"""
import ctypes
import itertools

import numpy as np
import pytest
import torch

from lbmpy_amd import LBStencil, Stencil
from lbmpy_amd import _native as nat
from lbmpy_amd.advanced_streaming import (LBMPeriodicityHandling, GpuArrays, Timestep, get_communication_slices,
                                          get_timesteps, streaming_patterns)
from lbmpy_amd.advanced_streaming.utility import access_table, kernel_code
from lbmpy_amd.slicing import get_ghost_region_slice, get_slice_before_ghost_layer
from oracle.lbm_oracle import streaming_accessor

STENCILS = ["D2Q9", "D3Q19", "D3Q27"]
KERNELS = [("pull", 0), ("push", 0), ("aa", 0), ("aa", 1), ("esotwist", 0), ("esotwist", 1), ("esopull", 0),
           ("esopull", 1), ("esopush", 0), ("esopush", 1)]


def _st(name):
    return LBStencil(Stencil[name])


def _oracle_tables(st, pattern, t):
    c = np.array(st.stencil_entries)
    return streaming_accessor(c, np.array(st.inverse_indices()), pattern, t)


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: f"{k[0]}-{k[1]}")
@pytest.mark.parametrize("stencil", STENCILS)
def test_kernel_access_tables_match_oracle(stencil, kernel):
    st = _st(stencil)
    pattern, t = kernel
    tab = access_table(st, kernel_code(pattern, Timestep(t)))
    roff, rslot, woff, wslot = _oracle_tables(st, pattern, t)
    D = st.D
    np.testing.assert_array_equal(tab[:, :D], roff)
    np.testing.assert_array_equal(tab[:, 3], rslot)
    np.testing.assert_array_equal(tab[:, 4:4 + D], woff)
    np.testing.assert_array_equal(tab[:, 7], wslot)
    assert not tab[:, D:3].any() and not tab[:, 4 + D:7].any()


@pytest.mark.parametrize("timestep", [Timestep.EVEN, Timestep.ODD])
@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", STENCILS)
def test_slices_not_empty_and_same_shape(stencil, pattern, timestep):
    st = _st(stencil)
    arr = np.zeros((4,) * st.D + (st.Q,))
    slices = get_communication_slices(st, streaming_pattern=pattern, prev_timestep=timestep, ghost_layers=1)
    for pairs in slices.values():
        for src, dst in pairs:
            assert all(n != 0 for n in arr[src].shape) and all(n != 0 for n in arr[dst].shape)
            assert arr[src].shape == arr[dst].shape


@pytest.mark.parametrize("stencil", STENCILS)
def test_pull_communication_slices(stencil):
    st = _st(stencil)
    slices = get_communication_slices(st, streaming_pattern="pull", prev_timestep=Timestep.BOTH, ghost_layers=1)
    from lbmpy_amd.advanced_streaming.communication import _length_one_to_int
    for i, d in enumerate(st):
        if i == 0:
            continue
        src, dst = next((s[0][:-1], s[1][:-1]) for s in slices[d] if s[0][-1] == i)
        assert src == _length_one_to_int(get_slice_before_ghost_layer(d, ghost_layers=1))
        assert dst == _length_one_to_int(get_ghost_region_slice(tuple(-e for e in d), ghost_layers=1))


def _cpu_pdf(shape, Q):
    """A CPU tensor with the PDF arrays' layout: storage [q][z][y][x], indexed [x, y, (z,) q]."""
    store = torch.zeros((Q,) + tuple(reversed(shape)), dtype=torch.float64)
    return store.permute(*range(len(shape), 0, -1), 0)


def _apply_boxes(arr, boxes):
    """CPU emulation of lbm355_copy_boxes on a tensor's storage."""
    flat = arr.as_strided((arr.untyped_storage().nbytes() // arr.element_size(),), (1,)).numpy()
    st = arr.stride()
    sz = st[2] if arr.dim() == 4 else 0
    for s_off, d_off, ex, ey, ez in boxes:
        i, j, k = np.meshgrid(np.arange(ex), np.arange(ey), np.arange(ez), indexing="ij")
        o = (i * st[0] + j * st[1] + k * sz).ravel()
        flat[d_off + o] = flat[s_off + o]


def _handler(st, shape, pattern, periodicity):
    arr = _cpu_pdf(tuple(n + 2 for n in shape), st.Q)
    h = LBMPeriodicityHandling(st, GpuArrays(periodicity, pdfs=arr), "pdfs", streaming_pattern=pattern)
    return h, arr


SHAPES = {2: (5, 4), 3: (5, 4, 3)}


def _check_closed_form(st, pattern, t, periodicity, drop=None):
    """Returns the number of (cell, direction) reads of the next kernel that differ from periodic addressing."""
    D, Q = st.D, st.Q
    shape = SHAPES[D]
    h, arr = _handler(st, shape, pattern, periodicity)
    prev = t if pattern not in ("pull", "push") else 0
    nxt = (1 - t) if pattern not in ("pull", "push") else 0
    _, _, woff, wslot = _oracle_tables(st, pattern, prev)
    roff, rslot, _, _ = _oracle_tables(st, pattern, nxt)
    rng = np.random.default_rng(11)
    f = rng.random((Q,) + shape)                          # post-collision value of direction i at cell x
    cells = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), axis=-1).reshape(-1, D)
    # periodic domain: the write and read addresses wrap on periodic axes
    per = np.full((Q,) + shape, np.nan)
    for i in range(Q):
        tgt = cells + woff[i]
        ok = np.ones(len(cells), bool)
        for a in range(D):
            if periodicity[a]:
                tgt[:, a] %= shape[a]
            else:
                ok &= (tgt[:, a] >= 0) & (tgt[:, a] < shape[a])
        per[(wslot[i],) + tuple(tgt[ok].T)] = f[(i,) + tuple(cells[ok].T)]
    # ghost-layer array: writes land in ghosts, then the copies
    arr.fill_(float("nan"))
    for i in range(Q):
        tgt = cells + woff[i] + 1
        arr[tuple(torch.from_numpy(tgt[:, a]) for a in range(D)) + (int(wslot[i]),)] = \
            torch.from_numpy(f[(i,) + tuple(cells.T)])
    boxes = h.boxes(arr, Timestep(prev) if pattern not in ("pull", "push") else Timestep.BOTH)
    if drop is not None:
        boxes = np.delete(boxes, drop, axis=0)
    _apply_boxes(arr, boxes)
    bad = 0
    for j in range(Q):
        src = cells + roff[j]
        ok = np.ones(len(cells), bool)
        for a in range(D):
            if periodicity[a]:
                continue
            ok &= (src[:, a] >= 0) & (src[:, a] < shape[a])    # reads across a non-periodic face: not compared
        wrapped = src.copy()
        for a in range(D):
            if periodicity[a]:
                wrapped[:, a] %= shape[a]
        want = per[(rslot[j],) + tuple(wrapped[ok].T)]
        got = arr[tuple(torch.from_numpy(src[ok][:, a] + 1) for a in range(D)) + (int(rslot[j]),)].numpy()
        if all(periodicity):
            assert not np.isnan(want).any()
        else:
            # next to a non-periodic face an in-place kernel also reads slots no interior cell writes (boundary
            # handling fills them): undefined on both sides
            keep = ~np.isnan(want)
            want, got = want[keep], got[keep]
        bad += int(np.sum(~(got == want)))
    return bad


@pytest.mark.parametrize("t", [0, 1])
@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", STENCILS)
def test_copies_reproduce_periodic_addressing(stencil, pattern, t):
    st = _st(stencil)
    if pattern in ("pull", "push") and t == 1:
        pytest.skip("one kernel")
    assert _check_closed_form(st, pattern, t, (True,) * st.D) == 0


@pytest.mark.parametrize("pattern", ["pull", "aa", "esotwist", "esopush"])
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q27"])
def test_copies_partly_periodic(stencil, pattern):
    st = _st(stencil)
    per = (True, False) if st.D == 2 else (True, False, True)
    for t in ((0,) if pattern == "pull" else (0, 1)):
        assert _check_closed_form(st, pattern, t, per) == 0


def test_planted_mistake_dropped_copy_is_detected():
    st = _st("D3Q19")
    assert _check_closed_form(st, "esopull", 1, (True,) * 3) == 0
    assert _check_closed_form(st, "esopull", 1, (True,) * 3, drop=5) > 0


@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", STENCILS)
def test_boxes_copy_what_the_slices_copy_and_are_disjoint(stencil, pattern):
    st = _st(stencil)
    shape = tuple(n + 2 for n in SHAPES[st.D])
    h, arr = _handler(st, SHAPES[st.D], pattern, (True,) * st.D)
    rng = np.random.default_rng(3)
    for t in get_timesteps(pattern):
        base = torch.from_numpy(rng.random(tuple(arr.shape)))
        arr.copy_(base)
        ref = base.clone()
        for src, dst in h.comm_slices[t.idx]:
            ref[dst] = ref[src]                                  # lbmpy's sequential slice copies
        boxes = h.boxes(arr, t)
        _apply_boxes(arr, boxes)
        assert torch.equal(arr, ref)
        st_ = arr.stride()
        sz = st_[2] if arr.dim() == 4 else 0
        srcs, dsts = [], []
        for s_off, d_off, ex, ey, ez in boxes:
            i, j, k = np.meshgrid(np.arange(ex), np.arange(ey), np.arange(ez), indexing="ij")
            o = (i * st_[0] + j * st_[1] + k * sz).ravel()
            srcs.append(s_off + o)
            dsts.append(d_off + o)
        srcs, dsts = np.concatenate(srcs), np.concatenate(dsts)
        assert len(np.unique(dsts)) == len(dsts)
        assert len(np.intersect1d(srcs, dsts)) == 0
        assert len(boxes) == len(h.comm_slices[t.idx])
    assert shape == tuple(arr.shape[:-1])


def test_copy_boxes_rejects_bad_boxes_without_a_gpu():
    lib = nat.load()
    good = np.array([[0, 10, 2, 2, 1]], dtype=np.int64)
    p = ctypes.c_void_p(4096)                 # never dereferenced: validation fails before any launch
    outside = np.array([[0, 95, 2, 2, 1]], dtype=np.int64)   # reaches 95 + 1 + 10 = 106 >= span 100
    empty = np.array([[0, 10, 0, 2, 1]], dtype=np.int64)
    for bad in (outside, empty):
        assert lib.lbm355_copy_boxes(nat.F64, p, 100, 1, 10, 0, bad.ctypes.data_as(ctypes.c_void_p), 1, None) \
            == 1
    assert lib.lbm355_copy_boxes(7, p, 100, 1, 10, 0, good.ctypes.data_as(ctypes.c_void_p), 1, None) == 1
    assert b"dtype" in lib.lbm355_last_error()
    assert lib.lbm355_copy_boxes(nat.F64, p, 100, 1, 10, 0, None, 0, None) == 0
    out = np.zeros(27 * 8, dtype=np.int32)
    assert lib.lbm355_pattern_access(5, 0, out.ctypes.data_as(ctypes.c_void_p)) < 0
    assert lib.lbm355_pattern_access(nat.D3Q27, 10, out.ctypes.data_as(ctypes.c_void_p)) < 0
    assert lib.lbm355_pattern_access(nat.D3Q27, nat.PUSH, out.ctypes.data_as(ctypes.c_void_p)) == 27


def test_handler_directions_follow_periodicity():
    st = _st("D3Q19")
    h, _ = _handler(st, (4, 4, 4), "pull", (True, False, True))
    for src, dst in h.comm_slices[0]:
        assert src[1] == slice(1, -1) and dst[1] == slice(1, -1)     # nothing crosses the y faces
    counts = {d: len(v) for d, v in get_communication_slices(st, comm_stencil=[(1, 0, 0), (1, 1, 0), (1, 1, 1)],
                                                              streaming_pattern="pull").items()}
    assert counts == {(1, 0, 0): 5, (1, 1, 0): 1, (1, 1, 1): 0}
    assert list(itertools.chain.from_iterable(get_communication_slices(
        st, comm_stencil=[(0, 0, 0)]).values())) == []
