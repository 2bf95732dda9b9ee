"""Race freedom of the in-place streaming kernels, shown the way a sanitizer would: every in-place pattern's
EVEN and ODD kernel, as the HIP sweeps are compiled with it (the lbm355_pattern_access tables), is run on the CPU
over a periodic domain cell by cell in a shuffled order, reading and writing one array in place.  The result must
equal, bit for bit, the run that reads every cell before writing any (what a GPU launch computes when no two cells
touch the same location), whatever the order -- so no cell's writes reach another cell's reads."""
import numpy as np
import pytest

from lbmpy_amd import LBStencil, Stencil, Timestep
from lbmpy_amd.advanced_streaming.utility import access_table, kernel_code


def _collide(f, rng_w):
    # a deterministic, non-linear per-cell map that mixes every direction (any order-dependent aliasing between
    # cells changes its input and therefore its output)
    return np.tanh(f @ rng_w) + 0.5 * f[::-1] * f.sum()


def _sweep(pdf, table, order, w):
    """In place, cell by cell in `order` (flat interior indices), with periodic wrap-around."""
    shape = pdf.shape[:-1]
    D = len(shape)
    out = pdf
    for flat in order:
        x = np.array(np.unravel_index(flat, shape))
        f = np.array([out[tuple((x + table[i, 0:D]) % shape)][table[i, 3]] for i in range(len(table))])
        g = _collide(f, w)
        for i in range(len(table)):
            out[tuple((x + table[i, 4:4 + D]) % shape)][table[i, 7]] = g[i]
    return out


def _two_phase(pdf, table, w):
    """Every cell reads from the unchanged array, then all writes land."""
    shape = pdf.shape[:-1]
    D = len(shape)
    src = pdf.copy()
    out = pdf.copy()
    for flat in range(int(np.prod(shape))):
        x = np.array(np.unravel_index(flat, shape))
        f = np.array([src[tuple((x + table[i, 0:D]) % shape)][table[i, 3]] for i in range(len(table))])
        g = _collide(f, w)
        for i in range(len(table)):
            out[tuple((x + table[i, 4:4 + D]) % shape)][table[i, 7]] = g[i]
    return out


@pytest.mark.parametrize("pattern", ["aa", "esotwist", "esopull", "esopush"])
@pytest.mark.parametrize("stencil,shape", [("D2Q9", (5, 4)), ("D3Q19", (4, 3, 5)), ("D3Q27", (3, 4, 3))])
def test_shuffled_cell_order_equals_two_phase_run(stencil, shape, pattern):
    st = LBStencil(Stencil[stencil])
    rng = np.random.default_rng(7)
    w = rng.normal(0, 0.3, (st.Q, st.Q))
    pdf = rng.normal(0, 0.1, tuple(shape) + (st.Q,))
    n = int(np.prod(shape))
    for t in (Timestep.EVEN, Timestep.ODD):
        table = np.asarray(access_table(st, kernel_code(pattern, t)))
        want = _two_phase(pdf, table, w)
        for seed in range(3):
            order = np.random.default_rng(seed).permutation(n)
            got = _sweep(pdf.copy(), table, order, w)
            np.testing.assert_array_equal(got, want)
        pdf = want                                   # the next kernel starts from this one's result


def test_an_aliasing_table_is_caught():
    # the check has teeth: pull's table read in place lets later cells read values earlier cells wrote
    st = LBStencil(Stencil.D2Q9)
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, (st.Q, st.Q))
    pdf = rng.normal(0, 0.1, (5, 4, st.Q))
    table = np.asarray(access_table(st, kernel_code("pull", Timestep.BOTH)))
    got = _sweep(pdf.copy(), table, np.random.default_rng(0).permutation(20), w)
    assert not np.array_equal(got, _two_phase(pdf, table, w))
