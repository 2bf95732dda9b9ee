"""Pins of the oracle's streaming accessors (oracle.lbm_oracle.streaming_accessor, OracleDomain with
'push', 'esotwist', 'esopull', 'esopush'), the way lbmpy's own tests pin its patterns
(tests/advanced_streaming/test_fully_periodic_flow.py: every pattern gives the same flow):

* every pattern gives bit-identical post-collision populations to pull after every step, on a periodic
  domain and in a channel with NoSlip walls and a moving UBB lid as boundary links (the collision sees the
  same values, so streaming alone could make a difference);
* the generic in-place path applied to AA equals the hand-written AA path;
* the in-place tables are race-free: a cell reads exactly the (offset, slot) set it writes, so no other cell
  touches it;
* a planted mistake (one wrong write slot in EsoTwist's ODD kernel) breaks the equality with pull.
"""
import numpy as np
import pytest

import oracle.lbm_oracle as O
from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from oracle.lbm_oracle import OracleDomain, streaming_accessor
from tests.lbm_helpers import seeded_state

PATTERNS = ["push", "esotwist", "esopull", "esopush"]
CASES = [
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6, force=(1e-4, 2e-5)), (7, 6)),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.8, 1.2], force=(1e-5, 0, 0)), (5, 4, 6)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.9, compressible=True), (4, 5, 3)),
]


def _method(st, meth, extra):
    return create_lb_method(LBMConfig(stencil=LBStencil(Stencil[st]), method=meth, **extra))


def _plane(shape_g, axis, index):
    f = np.zeros(shape_g, dtype=bool)
    sl = [slice(None)] * len(shape_g)
    sl[axis] = index
    f[tuple(sl)] = True
    return f


def _domains(m, shape, pattern, walls):
    D = m.dim
    if walls:
        gs = tuple(n + 2 for n in shape)
        bnd = [(_plane(gs, 1, 0), {"kind": "noslip"}),
               (_plane(gs, 1, -1), {"kind": "ubb", "velocity": (0.02,) + (0.0,) * (D - 1), "density": None})]
        faces = {a: ("ghost" if a == 1 else "periodic") for a in range(D)}
    else:
        bnd, faces = None, {a: "periodic" for a in range(D)}
    return [OracleDomain(m, shape, faces=faces, pattern=p, boundaries=bnd) for p in ("pull", pattern)]


@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_pattern_equals_pull_every_step(case, pattern, walls):
    st, meth, extra, shape = case
    m = _method(st, meth, extra)
    f0 = seeded_state(m, shape, seed=7, du=0.03, noise=1e-3)
    ref, dom = _domains(m, shape, pattern, walls)
    ref.set_pdfs_natural(f0)
    dom.set_pdfs_natural(f0)
    for _ in range(6):
        ref.step()
        dom.step()
        np.testing.assert_array_equal(dom.pdfs_in_natural_layout(), ref.pdfs_in_natural_layout())
    assert np.abs(ref.pdfs_in_natural_layout() - f0).max() > 1e-4


@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_generic_inplace_path_equals_hand_written_aa(walls):
    m = _method("D3Q19", Method.TRT, dict(relaxation_rates=[1.8, 1.2], force=(1e-5, 0, 0)))
    shape = (5, 4, 6)
    f0 = seeded_state(m, shape, seed=3, du=0.03, noise=1e-3)
    a, b = _domains(m, shape, "aa", walls)
    for d in (a, b):
        d.set_pdfs_natural(f0)
    for _ in range(5):
        a.step()
        b._generic_step()
        np.testing.assert_array_equal(a.pdfs_in_natural_layout(), b.pdfs_in_natural_layout())


@pytest.mark.parametrize("pattern", ["aa", "esotwist", "esopull", "esopush"])
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
def test_inplace_tables_are_race_free(stencil, pattern):
    st = LBStencil(Stencil[stencil])
    c = np.array(st.stencil_entries)
    inv = np.array(st.inverse_indices())
    for t in (0, 1):
        roff, rslot, woff, wslot = streaming_accessor(c, inv, pattern, t)
        reads = {(tuple(roff[i]), int(rslot[i])) for i in range(st.Q)}
        writes = {(tuple(woff[i]), int(wslot[i])) for i in range(st.Q)}
        assert reads == writes and len(reads) == st.Q
        # each slot is read (and written) at one offset only, so no two cells share a location
        assert sorted(rslot.tolist()) == list(range(st.Q)) and sorted(wslot.tolist()) == list(range(st.Q))


def test_planted_mistake_wrong_write_slot(monkeypatch):
    orig = O.streaming_accessor

    def broken(c, inv, pattern, timestep):
        r, rs, w, ws = orig(c, inv, pattern, timestep)
        if pattern == "esotwist" and timestep == 1:
            ws = ws.copy()
            ws[[1, 2]] = ws[[2, 1]]
        return r, rs, w, ws

    monkeypatch.setattr(O, "streaming_accessor", broken)
    st, meth, extra, shape = CASES[1]
    m = _method(st, meth, extra)
    f0 = seeded_state(m, shape, seed=7, du=0.03, noise=1e-3)
    ref, dom = _domains(m, shape, "esotwist", False)
    ref.set_pdfs_natural(f0)
    dom.set_pdfs_natural(f0)
    for _ in range(4):         # the wrong slot shows once the following (correct) kernel reads it
        ref.step()
        dom.step()
    assert np.abs(dom.pdfs_in_natural_layout() - ref.pdfs_in_natural_layout()).max() > 1e-6
