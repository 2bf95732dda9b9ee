"""Checkpoint and resume of a run's state on the GPU path.

lbmpy keeps no run state beyond the PDF arrays (its disk caches hold symbolic derivations only), so a run is
resumed from the populations.  With in-place streaming the state also includes which kernel runs next: an AA
field between an odd and an even step stores each population in its neighbour's slot, so resuming with the wrong
parity scrambles the field.  With FP16 stochastic rounding the random stream of a launch depends on the kernel's
launch count, which is state too.  A checkpoint therefore holds:

* the PDF tensor (ghost layers included), in its storage type;
* the number of time steps run, and which AA kernel runs next (by default the parity of that number);
* each kernel's launch count (stochastic rounding);
* a description of the layout (stencil, shape, dtype, streaming pattern) that resuming checks.

Files are written with ``torch.save`` as a dict of CPU tensors, integers and strings, and read back with
``torch.load(..., weights_only=True)``, which executes nothing from the file."""
import torch

FORMAT = "lbmpy_amd-checkpoint-1"


def _layout(pdf, stencil_name, pattern):
    return {"stencil": str(stencil_name), "shape": list(pdf.shape), "stride": list(pdf.stride()),
            "dtype": str(pdf.dtype).replace("torch.", ""), "pattern": str(pattern)}


def save_pdf_state(path, pdf, *, time_step, stencil_name, pattern, kernels=(), next_odd=None):
    """Write the state of a run driven by hand (create_lb_function kernels): the PDF tensor `pdf`, the number of
    steps run so far `time_step`, whether the next in-place step is an odd one (`next_odd`, by default
    `time_step` even) and the launch counts of `kernels`."""
    if next_odd is None:
        next_odd = time_step % 2 == 0
    state = {"format": FORMAT, "layout": _layout(pdf, stencil_name, pattern), "time_step": int(time_step),
             "next_odd": int(bool(next_odd)),
             "rounding_launches": [int(getattr(k, "rounding_launches", 0)) for k in kernels],
             # the storage order of the tensor is kept: the kernels address it through its strides
             "pdf": pdf.detach().to("cpu").clone()}
    torch.save(state, path)


def load_pdf_state(path, pdf, *, stencil_name, pattern, kernels=()):
    """Restore a state written by save_pdf_state into the existing tensor `pdf` (same stencil, shape, strides,
    dtype and streaming pattern) and the launch counts of `kernels`; returns (steps run, next step odd)."""
    state = torch.load(path, map_location="cpu", weights_only=True)
    if not isinstance(state, dict) or state.get("format") != FORMAT:
        raise ValueError(f"{path} is not a {FORMAT} file")
    want = _layout(pdf, stencil_name, pattern)
    if state["layout"] != want:
        raise ValueError(f"checkpoint layout {state['layout']} does not match this run's {want}")
    launches = state["rounding_launches"]
    if kernels and len(launches) != len(kernels):
        raise ValueError(f"checkpoint holds launch counts of {len(launches)} kernels, this run has {len(kernels)}")
    pdf.copy_(state["pdf"].to(pdf.device))
    for k, n in zip(kernels, launches):
        if hasattr(k, "rounding_launches"):
            k.rounding_launches = int(n)
    return int(state["time_step"]), bool(state["next_odd"])
