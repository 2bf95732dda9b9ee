"""Pins of the oracle's D3Q27 cumulant corrections (oracle.lbm_oracle.OracleCollision with ``galilean`` /
``fourth_order``), by linear (von Neumann) analysis of the oracle's own collision.

J = d(collision)/df at a uniform equilibrium (rho = 1, u = U), by central differences; for pull streaming on a
periodic lattice a Fourier mode exp(i k.x) advances by A(k) = diag(exp(-i k.c_q)) J per step.  The eigenvalue
of a hydrodynamic mode gives its damping per step, -ln|lambda| = Gamma k^2.

* Galilean correction (galilean_correction.py:21-91): the damping of a sound wave must not depend on a mean
  flow U along its wave vector.  Without the correction it changes by O(U^2) (the D3Q27 lattice's missing
  c^3 moments); with it the change must be at least ten times smaller.
* Fourth-order correction (fourth_order_correction.py:21-237): the shear viscosity measured from a shear wave
  whose strain has no normal components (k along an axis, or out-of-plane polarisation) must converge at
  fourth order in k at omega = 1.95 (error ratio > 11 per halving of k, against 4 without the correction) and
  be at least 20 times smaller than without it.
* The a-terms of the fourth-order correction act only on normal strain (an in-plane polarised diagonal
  shear wave).  As lbmpy writes them (prefactors 2/3, 2/3, -4/3) the leading error of that wave vanishes at
  twice their value, for every rate pair tried (measured 1.97-2.03); the test pins that ratio, so a dropped,
  sign-flipped or rescaled a-term fails it.
* The limiter (fourth_order_correction.py:185-209): with a vanishing limiter the limited third-order groups
  relax with rate 1, with a huge one with the optimal rate; all three groups take the rate computed from the
  x components (c_xyy, c_xzz, c_xyz), so a state without them keeps the unlimited rate in every group.
Planted mistakes (sign of the Galilean terms, sign of b, unlimited rates left at 1) fail the pins.
"""
import numpy as np
import pytest

import oracle.lbm_oracle as O
from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from oracle.lbm_oracle import OracleCollision, equilibrium_pdfs, corrections_from_config


def _config(w, **kw):
    rates = {"relaxation_rates": [w, kw.pop("bulk")]} if "bulk" in kw else {"relaxation_rate": w}
    return LBMConfig(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, compressible=True, **rates, **kw)


def _operator(lc, U=(0.0, 0.0, 0.0)):
    m = create_lb_method(lc)
    coll = OracleCollision(m, {}, **corrections_from_config(lc))
    f0 = equilibrium_pdfs(m, np.ones((1, 1, 1)), np.array(U, float).reshape(3, 1, 1, 1)).reshape(27, 1)
    assert np.abs(coll(f0)[0] - f0).max() <= 1e-15          # the uniform equilibrium is a fixed point
    h = 1e-6
    P = np.repeat(f0, 54, axis=1)
    for j in range(27):
        P[j, 2 * j] += h
        P[j, 2 * j + 1] -= h
    out = coll(P)[0]
    return m, (out[:, 0::2] - out[:, 1::2]) / (2 * h)


def _modes(m, J, kvec):
    c = np.array(m.stencil.stencil_entries, float)
    lam, V = np.linalg.eig(np.diag(np.exp(-1j * c @ np.asarray(kvec, float))) @ J)
    weight = (np.abs(V) ** 2).sum(0)
    return lam, (c.T @ V), weight


def _damping(lam, k):
    return -np.log(np.abs(lam)) / k ** 2


def _shear_error(lc, L, kdir, pol, w):
    """Relative error of the viscosity measured from a shear wave of wavelength L along kdir, polarised
    along pol: Gamma / nu - 1, nu = (1/w - 1/2) / 3."""
    m, J = _operator(lc)
    kd = np.asarray(kdir, float) / np.linalg.norm(kdir)
    p = np.asarray(pol, float) / np.linalg.norm(pol)
    k = 2 * np.pi / L
    lam, mom, weight = _modes(m, J, k * kd)
    sel = np.argmax(np.abs(p @ mom) ** 2 / weight)
    return _damping(lam[sel], k) / ((1 / w - 0.5) / 3) - 1


def _sound_change(lc, U, L=32):
    """Relative change of the two sound modes' damping (k along x) from U = 0 to the mean flow U."""
    out = []
    for vel in ((0.0, 0.0, 0.0), U):
        m, J = _operator(lc, vel)
        k = 2 * np.pi / L
        lam, mom, weight = _modes(m, J, (k, 0.0, 0.0))
        idx = np.argsort(-np.abs(mom[0]) ** 2 / weight)[:2]
        out.append(np.sort(_damping(lam[idx], k)))
    return out[1] / out[0] - 1


# ---------------------------------------------------------------------------------------------- Galilean

@pytest.mark.parametrize("w", [1.6, 1.9])
def test_galilean_correction_restores_invariance_of_sound_damping(w):
    for U in ((0.05, 0.0, 0.0), (0.1, 0.0, 0.0)):
        plain = _sound_change(_config(w), U)
        corrected = _sound_change(_config(w, galilean_correction=True), U)
        # the uncorrected scheme has the O(U^2) defect (about -4.6 U^2 / 0.1^2 percent)
        assert np.all(plain < -0.5 * U[0] ** 2) and np.all(plain > -8.0 * U[0] ** 2)
        assert np.abs(corrected).max() <= 0.1 * np.abs(plain).min()
    # a mean flow across the wave vector leaves the sound damping alone in both schemes
    assert np.abs(_sound_change(_config(w), (0.0, 0.1, 0.0))).max() <= 1e-8
    assert np.abs(_sound_change(_config(w, galilean_correction=True), (0.0, 0.1, 0.0))).max() <= 2e-4


def test_planted_mistake_galilean_sign(monkeypatch):
    orig = OracleCollision._add_galilean_correction

    def flipped(self, Cp, c, rho, u3, rates):
        before = Cp.copy()
        orig(self, Cp, c, rho, u3, rates)
        Cp[:] = 2 * before - Cp

    monkeypatch.setattr(OracleCollision, "_add_galilean_correction", flipped)
    U = (0.1, 0.0, 0.0)
    plain = _sound_change(_config(1.9), U)
    corrected = _sound_change(_config(1.9, galilean_correction=True), U)
    assert np.abs(corrected).max() > 0.1 * np.abs(plain).min()


# ---------------------------------------------------------------------------------------------- fourth order

AXIS_MODES = [((1, 0, 0), (0, 1, 0)), ((1, 1, 0), (0, 0, 1)), ((3, 0, 4), (0, 1, 0))]


def _fourth_order_pins(w):
    """Return (ok, details) of the shear-wave convergence pins at omega = w."""
    details = []
    ok = True
    for kdir, pol in AXIS_MODES:
        e32 = _shear_error(_config(w, fourth_order_correction=True), 32, kdir, pol, w)
        e64 = _shear_error(_config(w, fourth_order_correction=True), 64, kdir, pol, w)
        p64 = _shear_error(_config(w), 64, kdir, pol, w)
        details.append((kdir, e32, e64, p64))
        ok &= abs(e32 / e64) > 11.0 and abs(e64) <= abs(p64) / 20.0
    return ok, details


def test_fourth_order_correction_gives_fourth_order_shear_diffusion():
    ok, details = _fourth_order_pins(1.95)
    assert ok, details
    # without the correction the same waves converge at second order
    for kdir, pol in AXIS_MODES:
        p32 = _shear_error(_config(1.95), 32, kdir, pol, 1.95)
        p64 = _shear_error(_config(1.95), 64, kdir, pol, 1.95)
        assert 3.5 < p32 / p64 < 4.5


def test_fourth_order_correction_reduces_error_at_lower_rate():
    """At omega = 1.8 a small second-order remainder is left (about 1.5% of the uncorrected coefficient)."""
    for kdir, pol in AXIS_MODES:
        c = _shear_error(_config(1.8, fourth_order_correction=True), 64, kdir, pol, 1.8)
        p = _shear_error(_config(1.8), 64, kdir, pol, 1.8)
        assert abs(c) <= abs(p) / 20.0


@pytest.mark.parametrize("mistake", ["b_sign", "unlimited_rates_one"])
def test_planted_mistakes_fourth_order(monkeypatch, mistake):
    if mistake == "b_sign":
        orig = O._fourth_order_ab
        monkeypatch.setattr(O, "_fourth_order_ab", lambda w1, w2: (orig(w1, w2)[0], -orig(w1, w2)[1]))
    else:
        monkeypatch.setattr(O, "_fourth_order_omegas", lambda w1, w2: (1.0, 1.0, 1.0))
    ok, _ = _fourth_order_pins(1.95)
    assert not ok


@pytest.mark.parametrize("w1,w2", [(1.8, 1.0), (1.9, 1.2)])
def test_fourth_order_a_terms_cancel_normal_strain_error_at_twice_their_value(monkeypatch, w1, w2):
    orig = O._fourth_order_ab
    errs = {}
    for s in (0.0, 1.0):
        monkeypatch.setattr(O, "_fourth_order_ab", lambda a1, a2, s=s: (s * orig(a1, a2)[0], orig(a1, a2)[1]))
        errs[s] = _shear_error(_config(w1, bulk=w2, fourth_order_correction=True), 64, (1, 1, 0), (1, -1, 0), w1)
    zero = -errs[0.0] / (errs[1.0] - errs[0.0])
    assert abs(zero - 2.0) <= 0.06


def _third_order_groups(coll, f):
    """Polynomial cumulants x(y2+z2), y(x2+z2), z(x2+y2) and the x-group of a state (oracle transforms)."""
    rho, drho, u, _ = coll.macroscopic(f)
    u3 = [u[a] for a in range(3)]
    kappa = coll._shift(coll.Mmono @ f + coll.bg[:, None], [-x for x in u3])
    c = coll._cumulants_from_central(kappa, rho, u3)
    g = lambda e: c[coll.midx[e]]
    return np.stack([g((1, 2, 0)) + g((1, 0, 2)), g((2, 1, 0)) + g((0, 1, 2)), g((2, 0, 1)) + g((0, 2, 1))])


@pytest.mark.parametrize("limiter,expect", [(1e-12, "one"), (1e12, "optimal")])
def test_limiter_limits(limiter, expect):
    lc = _config(1.9, fourth_order_correction=limiter)
    m = create_lb_method(lc)
    coll = OracleCollision(m, {}, **corrections_from_config(lc))
    rng = np.random.default_rng(5)
    f = equilibrium_pdfs(m, np.ones((1, 1, 4)), 0.03 * rng.standard_normal((3, 1, 1, 4))).reshape(27, 4)
    f = f + 2e-3 * rng.standard_normal(f.shape)
    pre = _third_order_groups(coll, f)
    post = _third_order_groups(coll, coll(f)[0])
    w3 = O._fourth_order_omegas(1.9, 1.0)[0]
    want = np.zeros_like(pre) if expect == "one" else (1 - w3) * pre
    np.testing.assert_allclose(post, want, atol=1e-9 * np.abs(pre).max())


def test_limiter_uses_the_x_components_for_every_group():
    """A state mirror-symmetric in x has no c_xyy, c_xzz, c_xyz: with a vanishing limiter every group then keeps
    the unlimited optimal rate, also the y and z groups whose own cumulants (about 1e-3) are far above the
    limiter (1e-8; round-off leaves about 1e-17 in the x components)."""
    lc = _config(1.9, fourth_order_correction=1e-8)
    m = create_lb_method(lc)
    coll = OracleCollision(m, {}, **corrections_from_config(lc))
    rng = np.random.default_rng(9)
    u = 0.03 * rng.standard_normal((3, 1, 1, 4))
    u[0] = 0.0
    f = equilibrium_pdfs(m, np.ones((1, 1, 4)), u).reshape(27, 4) + 2e-3 * rng.standard_normal((27, 4))
    c = np.array(m.stencil.stencil_entries)
    mirror = [next(j for j in range(27) if tuple(c[j]) == (-c[i][0], c[i][1], c[i][2])) for i in range(27)]
    f = 0.5 * (f + f[mirror])
    pre = _third_order_groups(coll, f)
    assert np.abs(pre[0]).max() <= 1e-15 and np.abs(pre[1:]).min() > 1e-6
    post = _third_order_groups(coll, coll(f)[0])
    w3 = O._fourth_order_omegas(1.9, 1.0)[0]
    np.testing.assert_allclose(post[1:], (1 - w3) * pre[1:], rtol=1e-8)
