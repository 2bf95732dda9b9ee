"""What the trained closure gains: decaying D3Q27 cumulant turbulence (k0 = 4, u_rms = 0.04, omega = 1.99) on a
64^3 coarse grid against a 256^3 fine run box-filtered by 4, 400 coarse steps.  The network is trained on the
fine run of seed 0 (GPU kernels), its output scale and offset are calibrated on that run, and all models are
judged on a new flow (seed 1): the learned closure's kinetic-energy history must be closer to the filtered
fine run than both the method without a subgrid model and the Smagorinsky fallback (Cs = 0.12), and its final
spectrum no further off than either."""
import pytest

pytestmark = pytest.mark.gpu


def test_learned_closure_beats_no_model_and_smagorinsky():
    from lbmpy_amd.closure import turbulence_study
    print("closure study 64^3 / 256^3", flush=True)
    r = turbulence_study(n_coarse=64, factor=4, steps=400, snap_every=50, train_from=100, epochs=10, batch=32768,
                         seed=0, log=lambda s: print(s, flush=True))
    m = r["models"]
    ke = {k: v["ke_rel_l2_err"] for k, v in m.items()}
    sp = {k: v["spectrum_log_err"] for k, v in m.items()}
    print("KE errors", ke, "spectrum errors", sp, "alpha", r["alpha"], "beta", r["beta"], flush=True)
    assert ke["learned"] < ke["none"]
    assert ke["learned"] < ke["smagorinsky"]
    assert sp["learned"] <= sp["none"]
    assert sp["learned"] <= sp["smagorinsky"]
