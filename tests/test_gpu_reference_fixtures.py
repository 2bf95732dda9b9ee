"""The GPU path against lbmpy's own result files: the twelve D2Q9 SRT runs of tests/reference_fixtures.py
(force-driven channels and lid-driven cavities with the Guo, Luo and Simple force models, absolute and
zero-centred storage), set up through the scenario API as lbmpy's scenarios.py builds them and run by
``LatticeBoltzmannStep`` on the HIP sweep, the fused NoSlip faces and the UBB link kernel.

PDFs, density and velocity (the GPU macroscopic getter) must match the files to max-abs 1e-12 (FP64).
"""
import numpy as np
import pytest

from lbmpy_amd.scenarios import create_channel, create_lid_driven_cavity
from tests import reference_fixtures as rf

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("name", rf.fixture_names())
def test_gpu_reproduces_lbmpy_result_file(name):
    kw, run = rf.setup(name)
    if run["kind"] == "channel":
        # duct=True: the 2-D pipe wall of a full-height channel flags no cell, so both variants are this run
        kw.pop("force")
        step = create_channel(domain_size=run["shape"], force=run["force"][0], duct=True, **kw)
    else:
        step = create_lid_driven_cavity(domain_size=run["shape"], lid_velocity=run["lid"], **kw)
    # the step starts as lbmpy's does: rho = 1, u = 0 through the macroscopic setter
    step.run(run["steps"])
    ref = rf.load(name)
    pdfs = step.pdf_array[1:-1, 1:-1].cpu().numpy()
    assert np.abs(pdfs - ref["pdfs"]).max() <= TOL
    assert np.abs(step.density[:, :] - ref["rho"]).max() <= TOL
    assert np.abs(step.velocity[:, :] - ref["vel"]).max() <= TOL
