"""Golden vectors for the 64^3 Taylor-Green vortex (D3Q27 cumulant, AA, zero-centred, fully periodic,
u0 = 0.05, Re = 1600 with L = N / 2 pi), written by the C oracle after 500 and 2000 steps.

The full state is 56 MB, so the fixture keeps what pins it compactly:
  * PDFs of every 8th cell along each axis (8^3 cells x 27), in natural (pull) order;
  * per z-plane, per population sums over the plane (64 x 27), which any local error moves.

Run from the repository root (takes minutes on 8 cores):  python tests/golden/make_tgv64.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N = 64
U0 = 0.05
SNAPSHOTS = (500, 2000)
SAMPLE = 8
OUT = os.path.join(HERE, "tgv64_cumulant_aa.npz")


def tgv_method():
    from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
    nu = U0 * (N / (2 * np.pi)) / 1600.0
    omega = 1.0 / (3.0 * nu + 0.5)
    return create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT,
                                      relaxation_rate=omega, compressible=True, zero_centered=True,
                                      streaming_pattern="aa")), omega


def tgv_initial_pdfs(method):
    """Equilibrium of the TGV (rho from the pressure field, cell-centred coordinates), shape (27, z, y, x)."""
    from oracle.lbm_oracle import equilibrium_pdfs
    k = 2 * np.pi / N
    c = (np.arange(N) + 0.5) * k
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    u = np.stack([U0 * np.sin(x) * np.cos(y) * np.cos(z), -U0 * np.cos(x) * np.sin(y) * np.cos(z), 0 * x])
    rho = 1.0 + 3.0 * (U0 ** 2 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    return equilibrium_pdfs(method, rho, u)


def summarise(f):
    """(sampled PDFs, plane sums) of a natural-order state (27, z, y, x)."""
    return f[:, ::SAMPLE, ::SAMPLE, ::SAMPLE].copy(), f.sum(axis=(2, 3)).T.copy()


def main():
    from oracle.c_oracle import COracle
    m, omega = tgv_method()
    f0 = tgv_initial_pdfs(m)
    inv = np.array(m.stencil.inverse_indices())
    a = np.ascontiguousarray(f0[inv])
    co = COracle(m)
    out = {}
    t0 = time.time()
    for t in range(max(SNAPSHOTS)):
        co.sweep(a, a, "aa_odd" if t % 2 == 0 else "aa_even", nthreads=os.cpu_count())
        if t + 1 in SNAPSHOTS:
            s, p = summarise(a[inv])
            out[f"sample_{t + 1}"] = s
            out[f"planes_{t + 1}"] = p
            print(f"step {t + 1}: {time.time() - t0:.0f} s", flush=True)
    meta = dict(n=N, u0=U0, omega=omega, snapshots=list(SNAPSHOTS), sample=SAMPLE, method="D3Q27 cumulant AA",
                init="oracle equilibrium_pdfs of the TGV")
    np.savez_compressed(OUT, meta=json.dumps(meta), **out)
    print(OUT)


if __name__ == "__main__":
    main()
