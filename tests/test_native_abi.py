"""The C-ABI library loads and exports every entry point that ``include/lbm355.h`` declares.

Runs on the CPU: only host-side helpers (descriptor size, table sizes, version) are called; no
kernel is launched.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm355.h")


def _header_text():
    with open(HEADER) as fh:
        return fh.read()


def _declared_functions():
    text = re.sub(r"/\*.*?\*/", "", _header_text(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(lbm355_[a-z0-9_]+)\s*\(", text)))


def _header_macros():
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(LBM355_[A-Z0-9_]+)\s+\(?(-?\d+)\)?",
                                                              _header_text())}


@pytest.fixture(scope="module")
def lib():
    from lbmpy_amd.csrc.build import library_path
    if not os.path.exists(library_path()):
        pytest.skip("liblbm355.so not built (run __graft_entry__.build())")
    from lbmpy_amd import _native
    return _native.load()


def test_every_declared_symbol_is_exported(lib):
    from lbmpy_amd import _native
    declared = _declared_functions()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/lbm355.h but not exported"
    assert sorted(_native.EXPORTED) == declared


def test_descriptor_layout_matches(lib):
    from lbmpy_amd import _native
    assert lib.lbm355_desc_size() == ctypes.sizeof(_native.SweepDesc) == 200
    assert lib.lbm355_version() >= 5


def test_header_constants_match_binding():
    from lbmpy_amd import _native as n
    macros = _header_macros()
    expect = {
        "LBM355_D2Q9": n.D2Q9, "LBM355_D3Q19": n.D3Q19, "LBM355_D3Q27": n.D3Q27,
        "LBM355_FAMILY_MOMENT": n.FAMILY_MOMENT, "LBM355_FAMILY_CENTRAL": n.FAMILY_CENTRAL,
        "LBM355_FAMILY_CUMULANT": n.FAMILY_CUMULANT,
        "LBM355_PULL": n.PULL, "LBM355_AA_EVEN": n.AA_EVEN, "LBM355_AA_ODD": n.AA_ODD,
        "LBM355_F64": n.F64, "LBM355_F32": n.F32, "LBM355_F16": n.F16,
        "LBM355_FACE_GHOST": n.FACE_GHOST, "LBM355_FACE_PERIODIC": n.FACE_PERIODIC,
        "LBM355_FACE_NOSLIP": n.FACE_NOSLIP,
        "LBM355_FLAG_ZERO_CENTERED": n.FLAG_ZERO_CENTERED, "LBM355_FLAG_COMPRESSIBLE": n.FLAG_COMPRESSIBLE,
        "LBM355_FLAG_EQ_SHIFT": n.FLAG_EQ_SHIFT, "LBM355_FLAG_FORCE": n.FLAG_FORCE,
        "LBM355_FLAG_SMAGORINSKY": n.FLAG_SMAGORINSKY, "LBM355_FLAG_LEARNED_CLOSURE": n.FLAG_LEARNED_CLOSURE,
        "LBM355_FLAG_GALILEAN": n.FLAG_GALILEAN, "LBM355_FLAG_FOURTH_ORDER": n.FLAG_FOURTH_ORDER,
        "LBM355_ESOTWIST_EVEN": n.ESOTWIST_EVEN, "LBM355_ESOTWIST_ODD": n.ESOTWIST_ODD,
        "LBM355_ESOPULL_EVEN": n.ESOPULL_EVEN, "LBM355_ESOPULL_ODD": n.ESOPULL_ODD,
        "LBM355_ESOPUSH_EVEN": n.ESOPUSH_EVEN, "LBM355_ESOPUSH_ODD": n.ESOPUSH_ODD, "LBM355_PUSH": n.PUSH,
        "LBM355_BC_NOSLIP": n.BC_NOSLIP, "LBM355_BC_UBB": n.BC_UBB, "LBM355_BC_FREESLIP": n.BC_FREESLIP,
        "LBM355_BC_FIXED_DENSITY": n.BC_FIXED_DENSITY, "LBM355_BC_OUTFLOW": n.BC_OUTFLOW,
        "LBM355_BC_OUTFLOW_INIT": n.BC_OUTFLOW_INIT,
    }
    for k, v in expect.items():
        assert macros.get(k) == v, k


@pytest.mark.parametrize("stencil,family", [("D2Q9", "moment"), ("D3Q19", "moment"), ("D3Q27", "moment"),
                                            ("D2Q9", "central"), ("D3Q19", "central"), ("D3Q27", "central"),
                                            ("D3Q27", "cumulant")])
def test_table_sizes_agree_with_host_tables(lib, stencil, family):
    """The device table the host builds has exactly the element count the kernels index."""
    from lbmpy_amd import _native, LBMConfig, LBStencil, Method, create_lb_method
    from lbmpy_amd.kernel_tables import method_table
    method = {"moment": Method.TRT, "central": Method.CENTRAL_MOMENT, "cumulant": Method.CUMULANT}[family]
    kw = {"relaxation_rates": [1.7, 1.1]} if method == Method.TRT else {"relaxation_rate": 1.7}
    m = create_lb_method(LBMConfig(stencil=LBStencil(stencil), method=method, compressible=True, **kw))
    tab, _ = method_table(m, {})
    fam = {"moment": _native.FAMILY_MOMENT, "central": _native.FAMILY_CENTRAL,
           "cumulant": _native.FAMILY_CUMULANT}[family]
    assert lib.lbm355_table_elements(_native.STENCIL_NAMES[stencil], fam) == tab.size


def test_invalid_table_query_reports_error(lib):
    from lbmpy_amd import _native
    assert lib.lbm355_table_elements(_native.D3Q19, _native.FAMILY_CUMULANT) <= 0
    assert lib.lbm355_last_error()


# ---- argument validation: every case below is rejected before any HIP call, so it runs without a GPU

DUMMY_A, DUMMY_B = ctypes.c_void_p(0x10000), ctypes.c_void_p(0x20000)     # never dereferenced


def _desc(stencil="D3Q27", family="cumulant", pattern="AA_EVEN", faces=("PERIODIC",) * 3, flags=0):
    from lbmpy_amd import _native as n
    d = n.SweepDesc()
    d.stencil = n.STENCIL_NAMES[stencil]
    d.family = {"moment": n.FAMILY_MOMENT, "central": n.FAMILY_CENTRAL, "cumulant": n.FAMILY_CUMULANT}[family]
    d.pattern = getattr(n, pattern) if isinstance(pattern, str) else pattern
    d.store_dtype = d.compute_dtype = n.F32
    d.nx, d.ny, d.nz = 64, 8, 8
    d.ghost_layers = 1
    for a in range(3):
        d.face[a] = getattr(n, "FACE_" + faces[a])
    d.stride_y, d.stride_z = 66, 66 * 10
    d.stride_q = 66 * 10 * 10
    d.flags = flags
    d.tables = 0x30000
    return d


def _sweep_error(lib, d, src=DUMMY_A, dst=None):
    rc = lib.lbm355_sweep(ctypes.byref(d), src, src if dst is None else dst, None, None, None)
    return rc, lib.lbm355_last_error().decode()


def test_sweep_rejects_bad_patterns_and_array_counts(lib):
    assert _sweep_error(lib, _desc(pattern=42))[0] != 0
    rc, msg = _sweep_error(lib, _desc(pattern="ESOTWIST_EVEN", faces=("PERIODIC", "NOSLIP", "PERIODIC")))
    assert rc != 0 and "boundary links" in msg
    rc, msg = _sweep_error(lib, _desc(pattern="PULL"))
    assert rc != 0 and "two arrays" in msg
    rc, msg = _sweep_error(lib, _desc(pattern="PUSH"))
    assert rc != 0 and "two arrays" in msg
    rc, msg = _sweep_error(lib, _desc(pattern="ESOPULL_ODD"), DUMMY_A, DUMMY_B)
    assert rc != 0 and "in-place" in msg


def test_sweep_rejects_bad_correction_combinations(lib):
    from lbmpy_amd import _native as n
    rc, msg = _sweep_error(lib, _desc(stencil="D3Q19", family="moment", flags=n.FLAG_GALILEAN))
    assert rc != 0 and "cumulant" in msg
    rc, msg = _sweep_error(lib, _desc(flags=n.FLAG_FOURTH_ORDER | n.FLAG_SMAGORINSKY))
    assert rc != 0 and "Smagorinsky" in msg
    rc, msg = _sweep_error(lib, _desc(flags=n.FLAG_GALILEAN | n.FLAG_LEARNED_CLOSURE))
    assert rc != 0
    rc, msg = _sweep_error(lib, _desc(pattern="ESOPUSH_EVEN", flags=n.FLAG_LEARNED_CLOSURE))
    assert rc != 0 and "pull or AA" in msg


def test_sweep_rejects_unusable_cell_fields(lib):
    from lbmpy_amd import _native as n
    d = _desc()
    d.omega_out = 0x40000
    rc, msg = _sweep_error(lib, d)
    assert rc != 0 and "Smagorinsky" in msg
    d = _desc(stencil="D3Q19", family="moment")
    d.density_in = 0x40000
    rc, msg = _sweep_error(lib, d)
    assert rc != 0 and "velocity input" in msg
    d = _desc()
    d.velocity_in = 0x40000
    rc, msg = _sweep_error(lib, d)
    assert rc != 0 and "moment family" in msg
    d = _desc(stencil="D3Q19", family="moment")
    d.velocity_in = 0x40000
    d.out_stride_y, d.out_stride_z, d.out_stride_c = 66, 66 * 10, 66 * 10 * 10
    d.out_stride_y = 65                                          # a row of 64 cells and 2 ghost cells needs 66
    rc, msg = _sweep_error(lib, d)
    assert rc != 0 and "out_stride_y" in msg
    d.out_stride_y, d.out_stride_c = 66, 66 * 10 * 10 - 1
    rc, msg = _sweep_error(lib, d)
    assert rc != 0 and "out_stride_c" in msg


def test_boundary_links_reject_bad_patterns(lib):
    from lbmpy_amd import _native as n
    b = n.BoundaryDesc()
    b.stencil, b.store_dtype, b.kind = n.D2Q9, n.F64, n.BC_NOSLIP
    b.nx, b.ny, b.nz, b.ghost_layers = 8, 8, 1, 1
    b.stride_y, b.stride_q = 10, 100
    b.pattern = 42
    assert lib.lbm355_boundary_links(ctypes.byref(b), DUMMY_A, DUMMY_B, 4, None, None) != 0
    b.pattern = n.ESOTWIST_ODD
    b.fused_noslip[1] = 1
    assert lib.lbm355_boundary_links(ctypes.byref(b), DUMMY_A, DUMMY_B, 4, None, None) != 0
    assert "fused NoSlip" in lib.lbm355_last_error().decode()


def test_macroscopic_kernels_reject_bad_patterns(lib):
    from lbmpy_amd import _native as n
    d = _desc(stencil="D2Q9", family="moment", faces=("PERIODIC", "NOSLIP", "PERIODIC"))
    d.nz, d.stride_z, d.stride_q = 1, 0, 660
    assert lib.lbm355_set_equilibrium(ctypes.byref(d), DUMMY_A, DUMMY_B, DUMMY_B, DUMMY_B, n.AA_ODD, None) != 0
    assert lib.lbm355_set_equilibrium(ctypes.byref(d), DUMMY_A, DUMMY_B, DUMMY_B, DUMMY_B, n.ESOPULL_EVEN, None) != 0
    assert lib.lbm355_get_macroscopic(ctypes.byref(d), DUMMY_A, DUMMY_B, DUMMY_B, 42, None) != 0
    assert lib.lbm355_get_macroscopic(ctypes.byref(d), DUMMY_A, DUMMY_B, DUMMY_B, n.ESOTWIST_ODD, None) != 0


def test_batch_rejects_eso_and_corrected_members(lib):
    from lbmpy_amd import _native as n
    out = ctypes.c_void_p()
    for d in (_desc(pattern="ESOTWIST_EVEN"), _desc(flags=n.FLAG_GALILEAN)):
        descs = (n.SweepDesc * 1)(d)
        ptrs = (ctypes.c_void_p * 1)(DUMMY_A)
        assert lib.lbm355_batch_create(descs, ptrs, ptrs, 1, ctypes.byref(out)) != 0
