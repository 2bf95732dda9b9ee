"""Checkpoint files (lbmpy_amd.checkpoint): a PDF tensor, the step count, the next AA parity and the kernels'
launch counts survive a round trip bit for bit, and a state is refused for a run of another layout."""
import pytest
import torch

from lbmpy_amd.checkpoint import save_pdf_state, load_pdf_state


class _Kernel:
    def __init__(self, n):
        self.rounding_launches = n


def _field(dtype=torch.float16, shape=(5, 4, 3, 19)):
    g = torch.Generator().manual_seed(3)
    # fzyx storage viewed as [x, y, z, q], as create_pdf_array lays it out
    base = torch.randn(shape[::-1], generator=g).to(dtype)
    return base.permute(*range(len(shape) - 1, -1, -1))


def test_round_trip_is_bit_exact(tmp_path):
    pdf = _field()
    ks = [_Kernel(17), _Kernel(18)]
    save_pdf_state(tmp_path / "s.pt", pdf, time_step=35, stencil_name="D3Q19", pattern="aa", kernels=ks)
    want = pdf.clone()
    pdf.zero_()
    fresh = [_Kernel(0), _Kernel(0)]
    step, next_odd = load_pdf_state(tmp_path / "s.pt", pdf, stencil_name="D3Q19", pattern="aa", kernels=fresh)
    assert step == 35 and next_odd is False
    assert torch.equal(pdf.view(torch.int16), want.view(torch.int16))
    assert [k.rounding_launches for k in fresh] == [17, 18]


def test_explicit_parity_is_kept(tmp_path):
    pdf = _field(torch.float32)
    save_pdf_state(tmp_path / "s.pt", pdf, time_step=10, stencil_name="D3Q19", pattern="aa", next_odd=False)
    assert load_pdf_state(tmp_path / "s.pt", pdf, stencil_name="D3Q19", pattern="aa") == (10, False)


@pytest.mark.parametrize("change", ["dtype", "shape", "pattern", "stencil"])
def test_other_layout_is_refused(tmp_path, change):
    pdf = _field(torch.float32)
    save_pdf_state(tmp_path / "s.pt", pdf, time_step=2, stencil_name="D3Q19", pattern="pull")
    other = {"dtype": _field(torch.float64), "shape": _field(torch.float32, (6, 4, 3, 19))}.get(change, pdf)
    with pytest.raises(ValueError):
        load_pdf_state(tmp_path / "s.pt", other, stencil_name="D3Q27" if change == "stencil" else "D3Q19",
                       pattern="aa" if change == "pattern" else "pull")


def test_foreign_file_is_refused(tmp_path):
    torch.save({"weights": torch.zeros(3)}, tmp_path / "x.pt")
    with pytest.raises(ValueError):
        load_pdf_state(tmp_path / "x.pt", _field(), stencil_name="D3Q19", pattern="aa")
