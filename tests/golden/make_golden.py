"""Writes the golden vectors in this folder: seeded PDF states and the FP64 oracle's result
after a few steps, one ``.npz`` per case (``f0``, ``expected`` and a JSON ``meta`` string).

Run from the repository root:  python tests/golden/make_golden.py

The same seeded setup is re-run by ``tests/test_golden.py`` through the oracle (CPU) and through
the HIP sweep (GPU), so a change of either that moves a result shows up against the committed
numbers, not only against each other.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# name, stencil, method, LBMConfig keywords, interior shape (x, y, [z]), faces per axis, pattern, steps, seed
CASES = [
    ("d2q9_srt_pull_wall", "D2Q9", "SRT", dict(relaxation_rate=1.6), (10, 6), ("periodic", "noslip"), "pull", 12, 11),
    ("d2q9_trt_aa_force", "D2Q9", "TRT", dict(relaxation_rates=[1.9, 1.2], compressible=True, force=[1e-4, -2e-5]),
     (8, 6), ("periodic", "periodic"), "aa", 12, 12),
    ("d3q19_trt_pull_guo_channel", "D3Q19", "TRT",
     dict(relaxation_rates=[1.97, 1.6], compressible=True, force=[1e-5, 0.0, 0.0]), (6, 5, 4),
     ("periodic", "noslip", "noslip"), "pull", 10, 13),
    ("d3q19_mrt_aa_zero_centered", "D3Q19", "MRT", dict(relaxation_rate=1.8, force=[1e-5, 0.0, 0.0]), (6, 4, 4),
     ("periodic", "noslip", "periodic"), "aa", 10, 14),
    ("d3q27_cumulant_aa_tgv_like", "D3Q27", "CUMULANT", dict(relaxation_rate=1.9848, compressible=True), (6, 4, 4),
     ("periodic", "periodic", "periodic"), "aa", 10, 15),
    ("d3q27_cumulant_pull_force", "D3Q27", "CUMULANT",
     dict(relaxation_rate=1.8, compressible=True, force=[2e-5, -1e-5, 3e-5]), (5, 4, 4),
     ("periodic", "periodic", "periodic"), "pull", 10, 16),
    ("d3q27_central_moment_pull_wall", "D3Q27", "CENTRAL_MOMENT",
     dict(relaxation_rate=1.7, compressible=True, force=[1e-5, 0.0, 0.0]), (5, 4, 4),
     ("periodic", "noslip", "periodic"), "pull", 10, 17),
]


def build_method(meta):
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
    kw = dict(meta["config"])
    if "force" in kw:
        kw["force"] = tuple(kw["force"])
    return create_lb_method(LBMConfig(stencil=LBStencil(Stencil[meta["stencil"]]), method=Method[meta["method"]],
                                      streaming_pattern=meta["pattern"], **kw))


def lbm_config_kwargs(meta):
    from lbmpy_amd import LBStencil, Stencil, Method
    kw = dict(meta["config"])
    if "force" in kw:
        kw["force"] = tuple(kw["force"])
    return dict(stencil=LBStencil(Stencil[meta["stencil"]]), method=Method[meta["method"]],
                streaming_pattern=meta["pattern"], **kw)


def faces_of(meta):
    return {a: f for a, f in enumerate(meta["faces"])}


def load_case(path):
    with np.load(path, allow_pickle=False) as z:
        return json.loads(str(z["meta"])), z["f0"].copy(), z["expected"].copy()


def main():
    from tests.lbm_helpers import seeded_state, oracle_run
    for name, st, meth, cfg, shape, faces, pattern, steps, seed in CASES:
        meta = dict(name=name, stencil=st, method=meth, config=cfg, shape=list(shape), faces=list(faces),
                    pattern=pattern, steps=steps, seed=seed)
        m = build_method(meta)
        f0 = seeded_state(m, tuple(shape), seed=seed, du=0.03, noise=2e-4)
        ref, _ = oracle_run(m, tuple(shape), faces_of(meta), pattern, f0, steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), f0=f0, expected=ref, meta=json.dumps(meta))
        print(f"{name}: {f0.shape} max|f - f0| = {np.abs(ref - f0).max():.3e}")


if __name__ == "__main__":
    main()
