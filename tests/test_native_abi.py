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
    assert lib.lbm355_desc_size() == ctypes.sizeof(_native.SweepDesc) == 160
    assert lib.lbm355_version() >= 1


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
