"""Learned subgrid closure for the D3Q27 cumulant sweep.

A per-cell MLP (10 -> 32 -> 32 -> 1, tanh) maps a cell's non-equilibrium second moments to an eddy
viscosity nu_t = nu_scale * softplus(o) - nu_shift; the shear relaxation rate becomes
omega = 1 / (1/omega_0 + 3 nu_t) (relaxation time kept above 1/2).  The network is trained on the a priori
eddy viscosity of box-filtered fine runs; the scale of its output and the constant offset nu_shift (which
gives back part of the explicit viscosity where the scheme's own numerical dissipation already acts) are
calibrated a posteriori on the training flow.  It stands where lbmpy's Smagorinsky model stands
(``turbulence_models.py:25-84``, ``LBMConfig.subgrid_scale_model``); lbmpy has no learned closure.

Everything runs on the GPU through the C ABI (``include/lbm355.h``):
  * inference fused into the cumulant sweep (``SweepKernel.set_closure``), BF16 MFMA per wavefront;
  * ``lbm355_closure_features`` / ``lbm355_closure_forward`` for batches of samples;
  * training: ``lbm355_closure_train_step`` (forward, backward and weight gradients on BF16 MFMA with
    FP32 accumulation) and ``lbm355_closure_adam``.

Training data come from the method itself: a fine D3Q27 cumulant run, box-filtered onto the coarse
grid (:func:`subgrid_samples`).  Torch is used for device memory and for the data plumbing around the
kernels (filtering, FFT diagnostics), never for the network's arithmetic.
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as nat

IN, IN_PAD, H = 10, 16, 32
OFF_W1T, OFF_W2T, OFF_W3, OFF_B1, OFF_B2, OFF_B3, MASTER_SIZE = 0, 512, 1536, 1568, 1600, 1632, 1633
# byte offsets in the device parameter block (struct ClosureParams)
P_W1T, P_W2T, P_W3, P_B1, P_B2, P_B3, P_NU, P_SHIFT, P_SCALE, P_NUSHIFT, P_SIZE = 0, 1024, 3072, 3136, 3264, 3392, \
    3396, 3400, 3440, 3480, 3488


def _bf16_bits(a):
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class ClosureModel:
    """Network parameters: an FP32 master copy (training) and the packed device block the kernels read."""

    def __init__(self, device="cuda", seed=0, nu_scale=1e-3):
        self.lib = nat.load()
        if self.lib.lbm355_closure_params_size() != P_SIZE or self.lib.lbm355_closure_master_size() != MASTER_SIZE:
            raise ImportError("closure parameter layout mismatch between closure.h and lbmpy_amd/closure.py")
        self.device = torch.device(device)
        rng = np.random.default_rng(seed)
        master = np.zeros(MASTER_SIZE, dtype=np.float32)
        w1t = rng.uniform(-1, 1, (H, IN_PAD)) * math.sqrt(6.0 / (IN + H))
        w1t[:, IN:] = 0.0
        master[OFF_W1T:OFF_W2T] = w1t.ravel()
        master[OFF_W2T:OFF_W3] = (rng.uniform(-1, 1, (H, H)) * math.sqrt(6.0 / (2 * H))).ravel()
        master[OFF_W3:OFF_B1] = rng.uniform(-1, 1, H) * math.sqrt(6.0 / (H + 1))
        self.master = torch.tensor(master, device=self.device)
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.step = 0
        self.nu_scale = float(nu_scale)
        self.nu_shift = 0.0
        self.in_shift = np.zeros(IN, dtype=np.float32)
        self.in_scale = np.ones(IN, dtype=np.float32)
        self.params = torch.zeros(P_SIZE, dtype=torch.uint8, device=self.device)
        self.pack()

    # ------------------------------------------------------------------ parameter block
    def pack(self):
        """Write the device parameter block from the master copy (BF16 weights, FP32 biases)."""
        m = self.master.cpu().numpy()
        blk = np.zeros(P_SIZE, dtype=np.uint8)
        blk[P_W1T:P_W2T] = _bf16_bits(m[OFF_W1T:OFF_W2T]).view(np.uint8)
        blk[P_W2T:P_W3] = _bf16_bits(m[OFF_W2T:OFF_W3]).view(np.uint8)
        blk[P_W3:P_B1] = _bf16_bits(m[OFF_W3:OFF_B1]).view(np.uint8)
        blk[P_B1:P_B2] = m[OFF_B1:OFF_B2].astype(np.float32).view(np.uint8)
        blk[P_B2:P_B3] = m[OFF_B2:OFF_B3].astype(np.float32).view(np.uint8)
        blk[P_B3:P_NU] = np.float32(m[OFF_B3]).reshape(1).view(np.uint8)
        blk[P_NU:P_SHIFT] = np.float32(self.nu_scale).reshape(1).view(np.uint8)
        blk[P_SHIFT:P_SCALE] = self.in_shift.astype(np.float32).view(np.uint8)
        blk[P_SCALE:P_NUSHIFT] = self.in_scale.astype(np.float32).view(np.uint8)
        blk[P_NUSHIFT:P_NUSHIFT + 4] = np.float32(self.nu_shift).reshape(1).view(np.uint8)
        self.params.copy_(torch.from_numpy(blk))

    def set_normalisation(self, features):
        """Per-feature shift/scale from a sample of raw features (mean and 1/std)."""
        f = features.double()
        self.in_shift = f.mean(0).float().cpu().numpy()
        std = f.std(0).float().cpu().numpy()
        self.in_scale = (1.0 / np.maximum(std, 1e-30)).astype(np.float32)
        self.pack()

    def set_output_bias(self, targets):
        """Start the output at the mean target: softplus(b3) = mean(y)."""
        y = float(targets.double().mean().clamp_min(1e-6))
        b3 = math.log(math.expm1(y)) if y < 20 else y
        self.master[OFF_B3] = b3
        self.pack()

    def state(self):
        return {"master": self.master.cpu().numpy(), "in_shift": self.in_shift, "in_scale": self.in_scale,
                "nu_scale": self.nu_scale, "nu_shift": self.nu_shift}

    def load_state(self, st):
        self.master.copy_(torch.as_tensor(np.asarray(st["master"], dtype=np.float32), device=self.device))
        self.in_shift = np.asarray(st["in_shift"], dtype=np.float32)
        self.in_scale = np.asarray(st["in_scale"], dtype=np.float32)
        self.nu_scale = float(st["nu_scale"])
        self.nu_shift = float(st.get("nu_shift", 0.0))
        self.pack()

    # ------------------------------------------------------------------ kernels
    def features(self, pi_neq, rho):
        """Raw features (n, 10) from non-equilibrium second moments (n, 6) and density (n,)."""
        pi_neq = pi_neq.float().contiguous()
        rho = rho.float().contiguous()
        out = torch.empty((pi_neq.shape[0], IN), dtype=torch.float32, device=self.device)
        nat.check(self.lib.lbm355_closure_features(_ptr(pi_neq), _ptr(rho), pi_neq.shape[0], _ptr(out),
                                                   nat.current_stream_handle()))
        return out

    def forward(self, features):
        """Network outputs o (n,)."""
        x = features.float().contiguous()
        out = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        nat.check(self.lib.lbm355_closure_forward(_ptr(self.params), _ptr(x), x.shape[0], _ptr(out),
                                                  nat.current_stream_handle()))
        return out

    def eddy_viscosity(self, features):
        return self.nu_scale * torch.nn.functional.softplus(self.forward(features))

    def gradient(self, features, targets, blocks=1024):
        """(loss, grad) of L = mean((softplus(o) - y)^2) at the current parameters (GPU kernels)."""
        x = features.float().contiguous()
        y = targets.float().contiguous()
        grad = torch.empty(MASTER_SIZE, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        nat.check(self.lib.lbm355_closure_train_step(_ptr(self.params), _ptr(x), _ptr(y), x.shape[0], _ptr(grad),
                                                     _ptr(loss), blocks, nat.current_stream_handle()))
        return loss, grad

    def adam_step(self, grad, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        nat.check(self.lib.lbm355_closure_adam(_ptr(self.master), _ptr(grad), _ptr(self.m), _ptr(self.v), lr, beta1,
                                               beta2, eps, self.step, _ptr(self.params), nat.current_stream_handle()))

    def train(self, features, targets, epochs=4, batch=65536, lr=2e-3, seed=0, log=None):
        """Adam on minibatches of (features, targets = nu_t / nu_scale); returns the per-epoch mean loss."""
        n = features.shape[0]
        g = torch.Generator(device=self.device).manual_seed(seed)
        history = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g, device=self.device)
            tot = torch.zeros((), device=self.device, dtype=torch.float64)
            nb = 0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                loss, grad = self.gradient(features[idx], targets[idx])
                self.adam_step(grad, lr=lr)
                tot += loss[0].double()
                nb += 1
            history.append(float(tot / nb))
            if log:
                log(f"closure training epoch {ep}: loss {history[-1]:.4e}")
        return history


# ---------------------------------------------------------------------- flow fields and diagnostics

def solenoidal_field(n, k0=4.0, u_rms=0.04, seed=0, device="cuda"):
    """Random divergence-free velocity (3, n, n, n) [component, z, y, x] with energy spectrum
    E(k) ~ k^4 exp(-2 (k/k0)^2) (k in box wavenumbers), scaled to the given rms per component."""
    gen = torch.Generator(device="cpu").manual_seed(seed)
    k1 = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64)          # integer wavenumbers
    kr = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64)
    kz, ky, kx = torch.meshgrid(k1, k1, kr, indexing="ij")
    kk = torch.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    amp = torch.where(kk > 0, torch.sqrt(kk ** 4 * torch.exp(-2 * (kk / k0) ** 2) / (4 * math.pi * kk ** 2 + 1e-300)),
                      torch.zeros_like(kk))
    re = torch.randn((3,) + kk.shape, generator=gen, dtype=torch.float64).to(device)
    im = torch.randn((3,) + kk.shape, generator=gen, dtype=torch.float64).to(device)
    kx, ky, kz, kk, amp = (t.to(device) for t in (kx, ky, kz, kk, amp))
    c = torch.complex(re, im) * amp
    del re, im
    kv = torch.stack([kx, ky, kz])
    kdotc = (kv * c).sum(0)
    c = c - kv * kdotc / torch.where(kk > 0, kk ** 2, torch.ones_like(kk))            # project out k || c
    u = torch.stack([torch.fft.irfftn(c[a], s=(n, n, n)) for a in range(3)])
    u = u * (u_rms / torch.sqrt((u ** 2).mean(dim=(1, 2, 3)).mean()))
    return u


def box_filter(a, f):
    """Average over f^3 boxes of the last three axes."""
    s = a.shape
    n = s[-1] // f
    return a.reshape(*s[:-3], n, f, n, f, n, f).mean(dim=(-5, -3, -1))


def cumulant_equilibrium_aa(storage, rho, u, chunk=16):
    """Writes the zero-centred D3Q27 cumulant-method equilibrium of (rho, u) into an AA field's storage
    (27, z, y, x) in the layout before the first (ODD) step: slot inv(i) holds f_i.
    f_i = rho prod_a phi_{c_ia}(u_a), phi_0 = 2/3 - u^2, phi_(+-1) = (1/3 + u^2 +- u)/2."""
    from .stencils import LBStencil
    st = LBStencil("D3Q27")
    inv = st.inverse_indices()
    w1 = {0: 2.0 / 3.0, 1: 1.0 / 6.0, -1: 1.0 / 6.0}
    nz = storage.shape[1]
    for z0 in range(0, nz, chunk):
        r = rho[z0:z0 + chunk].double()
        uu = [u[a, z0:z0 + chunk].double() for a in range(3)]

        def phi(c, ua):
            return (2.0 / 3.0 - ua * ua) if c == 0 else 0.5 * (1.0 / 3.0 + ua * ua + c * ua)
        for i, cv in enumerate(st):
            f = r * phi(cv[0], uu[0]) * phi(cv[1], uu[1]) * phi(cv[2], uu[2]) - w1[cv[0]] * w1[cv[1]] * w1[cv[2]]
            storage[inv[i], z0:z0 + chunk] = f.to(storage.dtype)


def macroscopic_aa(storage, chunk=16):
    """(rho, u) in FP64 of a zero-centred D3Q27 AA field after an even number of steps
    (slot j holds f_inv(j), so the momentum sums carry -c_j)."""
    from .stencils import LBStencil
    c = torch.tensor([list(d) for d in LBStencil("D3Q27")], dtype=torch.float64, device=storage.device)
    nz = storage.shape[1]
    rho = torch.empty(storage.shape[1:], dtype=torch.float64, device=storage.device)
    u = torch.empty((3,) + tuple(storage.shape[1:]), dtype=torch.float64, device=storage.device)
    for z0 in range(0, nz, chunk):
        g = storage[:, z0:z0 + chunk].double()
        r = g.sum(0) + 1.0
        rho[z0:z0 + chunk] = r
        u[:, z0:z0 + chunk] = -torch.einsum("qa,q...->a...", c, g) / r
    return rho, u


def kinetic_energy(u):
    return 0.5 * float((u ** 2).sum(0).mean())


def energy_spectrum(u):
    """Shell-summed spectrum E(k), k = 0..n/2, normalised so that sum_k E(k) = (1/2) <|u|^2>."""
    n = u.shape[-1]
    uh = torch.fft.fftn(u, dim=(1, 2, 3)) / n ** 3
    e = 0.5 * (uh.real ** 2 + uh.imag ** 2).sum(0)
    k1 = torch.fft.fftfreq(n, d=1.0 / n, device=u.device)
    kz, ky, kx = torch.meshgrid(k1, k1, k1, indexing="ij")
    shell = torch.round(torch.sqrt(kx ** 2 + ky ** 2 + kz ** 2)).long().clamp(max=n // 2)
    spec = torch.zeros(n // 2 + 1, dtype=torch.float64, device=u.device)
    spec.scatter_add_(0, shell.ravel(), e.ravel().double())
    return spec.cpu().numpy()


def _ddx(a, axis):
    return 0.5 * (torch.roll(a, -1, dims=axis) - torch.roll(a, 1, dims=axis))


def _box3(a):
    """Periodic 3x3x3 average over the last three axes."""
    out = a
    for ax in (-1, -2, -3):
        out = (torch.roll(out, 1, ax) + out + torch.roll(out, -1, ax)) / 3.0
    return out


def subgrid_samples(rho_f, u_f, factor, omega0, model):
    """Training samples from a fine-grid state (rho_f, u_f on [z, y, x]) box-filtered by `factor` onto the
    coarse grid.  The subgrid stress tau_ij = <rho u_i u_j>/<rho> - ~u_i ~u_j (Favre filter) and the filtered
    strain rate S_ij give the least-squares eddy viscosity nu_t = max(0, -<tau^d : S> / (2 <S : S>)) (local 3^3
    averages); the coarse cell's pre-collision non-equilibrium moments are Pi_neq = -2 rho c_s^2 tau_eff S with
    tau_eff = 1/omega0 + 3 nu_t (Chapman-Enskog).  Returns (features (n, 10), targets nu_t / nu_scale (n,))."""
    rb = box_filter(rho_f, factor)
    ub = box_filter(rho_f[None] * u_f, factor) / rb[None]
    axes = {0: -1, 1: -2, 2: -3}            # velocity component a -> tensor axis of coordinate a
    S = {}
    for a in range(3):
        for b in range(a, 3):
            S[(a, b)] = 0.5 * (_ddx(ub[a], axes[b]) + _ddx(ub[b], axes[a]))
    tau = {}
    for a in range(3):
        for b in range(a, 3):
            tau[(a, b)] = box_filter(rho_f * u_f[a] * u_f[b], factor) / rb - ub[a] * ub[b]
    trt = (tau[(0, 0)] + tau[(1, 1)] + tau[(2, 2)]) / 3.0
    num = torch.zeros_like(rb)
    den = torch.zeros_like(rb)
    for (a, b), s in S.items():
        w = 1.0 if a == b else 2.0
        td = tau[(a, b)] - (trt if a == b else 0.0)
        num = num + w * td * s
        den = den + w * s * s
    nu_t = torch.clamp(-_box3(num) / (2.0 * _box3(den) + 1e-30), min=0.0)
    tau_eff = 1.0 / omega0 + 3.0 * nu_t
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    pi = torch.stack([-(2.0 / 3.0) * rb * tau_eff * S[k] for k in order], dim=-1).reshape(-1, 6)
    feats = model.features(pi.float(), rb.reshape(-1).float())
    return feats, (nu_t.reshape(-1) / model.nu_scale).float()


# ---------------------------------------------------------------------- decaying turbulence study

def _aa_kernels(omega, closure=None, smagorinsky=None):
    from . import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target, Timestep
    from .enums import SubgridScaleModel
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float32", compute_type="float32", ghost_layers=0)
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=omega, compressible=True,
              zero_centered=True, streaming_pattern="aa")
    if smagorinsky is not None:
        kw["subgrid_scale_model"] = (SubgridScaleModel.SMAGORINSKY, smagorinsky)
    ks = [create_lb_function(lbm_config=LBMConfig(**kw, timestep=t), config=cfg) for t in (Timestep.ODD, Timestep.EVEN)]
    for k in ks:
        k.set_faces(["periodic"] * 3)
        if closure is not None:
            k.set_closure(closure)
    return ks


def _aa_field(n, rho, u):
    from .sweep import create_pdf_array
    pdf = create_pdf_array((n, n, n), 27, ghost_layers=0, dtype="float32")
    cumulant_equilibrium_aa(pdf.permute(3, 2, 1, 0), rho, u)
    return pdf


def fine_reference(n_coarse, factor, steps, snap_every, k0, u_rms, omega_coarse, seed, model=None, train_from=0,
                   log=print):
    """Fine run of the decaying-turbulence setup (n_coarse * factor cells per side, viscosity scaled by
    `factor`, `factor` fine steps per coarse step), box-filtered onto the coarse grid every `snap_every` coarse
    steps.  Returns the filtered kinetic energies, the final filtered spectrum, the filtered initial field and,
    if `model` is given, training samples from the snapshots at coarse step >= train_from."""
    import time
    n_f = n_coarse * factor
    nu_c = (1.0 / omega_coarse - 0.5) / 3.0
    omega_f = 1.0 / (3.0 * factor * nu_c + 0.5)
    t0 = time.time()
    u_fine = solenoidal_field(n_f, k0, u_rms, seed)
    rho_fine = torch.ones(u_fine.shape[1:], dtype=torch.float64, device=u_fine.device)
    u0 = box_filter(u_fine, factor)
    fine = _aa_field(n_f, rho_fine, u_fine)
    del u_fine, rho_fine
    ks = _aa_kernels(omega_f)
    ke, ts, feats, targs = [], [], [], []
    spec = None
    done, ms = 0, 0.0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for t in range(0, steps + 1):
        if t % snap_every == 0:
            rho_f, u_f = macroscopic_aa(fine.permute(3, 2, 1, 0))
            ub = box_filter(rho_f[None] * u_f, factor) / box_filter(rho_f, factor)[None]
            ke.append(kinetic_energy(ub))
            ts.append(t)
            if t == steps:
                spec = energy_spectrum(ub)
            if model is not None and t >= train_from:
                f_, y_ = subgrid_samples(rho_f, u_f, factor, omega_coarse, model)
                feats.append(f_)
                targs.append(y_)
            del rho_f, u_f, ub
        if t == steps:
            break
        e0.record()
        for s_ in range(factor):
            ks[(done + s_) % 2](src=fine)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
        done += factor
    del fine
    torch.cuda.empty_cache()
    mlups = n_f ** 3 * done / (ms / 1e3) / 1e6
    log(f"fine run {n_f}^3 (seed {seed}): {done} steps, {mlups:.0f} MLUPS, {time.time() - t0:.1f} s wall")
    out = {"ke": ke, "t": ts, "spectrum": spec, "u0": u0, "mlups": mlups, "omega_fine": omega_f}
    if model is not None:
        out["X"] = torch.cat(feats)
        out["Y"] = torch.cat(targs)
    return out


def coarse_run(n, u0, omega, steps, snap_every, closure=None, smagorinsky=None):
    """Coarse D3Q27 cumulant run from (rho = 1, u0): kinetic energy every `snap_every` steps, final spectrum and
    the sweep rate (HIP events around each launch)."""
    kern = _aa_kernels(omega, closure=closure, smagorinsky=smagorinsky)
    f = _aa_field(n, torch.ones(u0.shape[1:], dtype=torch.float64, device=u0.device), u0)
    ke, spec, ms = [], None, 0.0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for t in range(0, steps + 1):
        if t % snap_every == 0:
            _, u_c = macroscopic_aa(f.permute(3, 2, 1, 0))
            ke.append(kinetic_energy(u_c))
            if t == steps:
                spec = energy_spectrum(u_c)
        if t == steps:
            break
        e0.record()
        kern[t % 2](src=f)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
    return np.array(ke), spec, n ** 3 * steps / (ms / 1e3) / 1e6


def errors(ke, spec, ref):
    """Relative L2 error of the kinetic-energy history and mean |log| ratio of the final spectrum for
    1 <= k <= n/4 against a filtered fine reference."""
    rk = np.asarray(ref["ke"])
    kmax = (len(spec) - 1) // 2
    return (float(np.sqrt(((ke - rk) ** 2).sum() / (rk ** 2).sum())),
            float(np.mean(np.abs(np.log(spec[1:kmax + 1] / ref["spectrum"][1:kmax + 1])))))


def turbulence_study(n_coarse=128, factor=4, steps=1200, snap_every=100, train_from=200, k0=4.0, u_rms=0.04,
                     omega_coarse=1.99, smagorinsky=0.12, seed=0, epochs=20, batch=65536, lr=2e-3,
                     alphas=(0.0, 0.25, 0.5, 1.0), betas=(0.0, 0.25, 0.5, 0.75), log=print):
    """Decaying homogeneous turbulence, D3Q27 cumulant, FP32, AA, periodic; learned closure vs Smagorinsky vs
    no subgrid model, judged against the box-filtered fine run.

    1. Training data: the fine run of seed `seed` (:func:`fine_reference`), subgrid samples from its
       snapshots (:func:`subgrid_samples`, a priori eddy viscosity of the filtered field).
    2. The network is trained on the GPU (own forward/backward kernels, Adam).
    3. Its output scale alpha and a constant offset beta nu_0 (nu_t -> alpha nu_t - beta nu_0, nu_0 the
       molecular viscosity) are calibrated a posteriori on the same training run: coarse runs for each pair,
       the one closest to the filtered kinetic energy is kept (alpha = beta = 0 is the method without a
       subgrid model).
    4. Judgement on a new flow, seed + 1: a second fine run is the reference for coarse runs without subgrid
       model, with Smagorinsky (Cs = smagorinsky), with the learned closure and, as an ablation, with the best
       constant offset alone (alpha = 0)."""
    import time
    nu_c = (1.0 / omega_coarse - 0.5) / 3.0
    model = ClosureModel(seed=seed, nu_scale=10.0 * nu_c)
    common = dict(n_coarse=n_coarse, factor=factor, steps=steps, snap_every=snap_every, k0=k0, u_rms=u_rms,
                  omega_coarse=omega_coarse, log=log)
    train = fine_reference(seed=seed, model=model, train_from=train_from, **common)
    X, Y = train.pop("X"), train.pop("Y")
    model.set_normalisation(X)
    model.set_output_bias(Y)
    t1 = time.time()
    history = model.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    nu_pred = model.eddy_viscosity(X[: min(len(X), 1 << 20)])
    log(f"trained on {len(X)} samples in {train_s:.1f} s; target nu_t mean {float(Y.mean()) * model.nu_scale:.3e}, "
        f"predicted mean {float(nu_pred.mean()):.3e}")
    del X, Y
    base_scale = model.nu_scale
    calib = {}
    for al in alphas:
        for be in betas:
            model.nu_scale = base_scale * al
            model.nu_shift = be * nu_c
            model.pack()
            ke, spec, _ = coarse_run(n_coarse, train["u0"], omega_coarse, steps, snap_every,
                                     closure=model if (al > 0 or be > 0) else None)
            calib[(al, be)] = errors(ke, spec, train)
            log(f"calibration alpha {al} beta {be}: KE error {calib[(al, be)][0]:.4f}, "
                f"spectrum log error {calib[(al, be)][1]:.4f}")
    alpha, beta = min(calib, key=lambda k: calib[k][0])
    # ablation: the best constant offset without the network (alpha = 0)
    beta0 = min((k for k in calib if k[0] == 0.0), key=lambda k: calib[k][0])[1]
    offset_only = ClosureModel(seed=seed, nu_scale=base_scale)
    offset_only.load_state(model.state())
    offset_only.nu_scale = 0.0
    offset_only.nu_shift = beta0 * nu_c
    offset_only.pack()
    model.nu_scale = base_scale * alpha
    model.nu_shift = beta * nu_c
    model.pack()
    use = model if (alpha > 0 or beta > 0) else None
    log(f"calibrated output scale alpha = {alpha}, offset beta = {beta} (offset-only ablation: beta = {beta0})")

    ref = fine_reference(seed=seed + 1, **common)
    results = {}
    for name, kw in (("none", {}), ("smagorinsky", {"smagorinsky": smagorinsky}),
                     ("offset_only", {"closure": offset_only if beta0 > 0 else None}), ("learned", {"closure": use})):
        ke, spec, mlups = coarse_run(n_coarse, ref["u0"], omega_coarse, steps, snap_every, **kw)
        ke_err, sp_err = errors(ke, spec, ref)
        kmax = n_coarse // 4
        results[name] = {"ke": ke.tolist(), "ke_rel_l2_err": ke_err, "spectrum_log_err": sp_err,
                         "spectrum": spec[: kmax + 1].tolist(), "mlups": mlups}
        log(f"coarse {name} (seed {seed + 1}): KE error {ke_err:.4f}, spectrum log error {sp_err:.4f}, "
            f"{mlups:.0f} MLUPS")
    return {"n_coarse": n_coarse, "n_fine": n_coarse * factor, "steps": steps, "omega_coarse": omega_coarse,
            "omega_fine": ref["omega_fine"], "ref_ke": ref["ke"], "ref_t": ref["t"],
            "ref_spectrum": ref["spectrum"][: n_coarse // 4 + 1].tolist(), "train_loss": history,
            "train_seconds": train_s, "fine_mlups": ref["mlups"], "alpha": alpha, "beta": beta,
            "offset_only_beta": beta0,
            "calibration": {f"{k[0]},{k[1]}": v for k, v in calib.items()}, "models": results, "closure": model}
