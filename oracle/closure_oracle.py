"""CPU FP64 oracle of the learned-closure network -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import this
module.  It restates, in numpy, the network that ``lbmpy_amd/csrc/closure.h`` runs on the MFMA units:

    x' = bf16((x - shift) * scale)                       (10 features, padded to 16 with zeros)
    h1 = tanh(W1t x' + b1),  a1 = bf16(h1)               (32 units)
    h2 = tanh(W2t a1 + b2)                               (32 units)
    o  = w3 . h2 + b3                                    (w3 in BF16)
    nu_t = nu_scale * softplus(o) - nu_shift,  omega = 1 / max(1/omega_0 + 3 nu_t, 0.5 + 3e-6)

and its gradient for the training loss L = (1/n) sum (softplus(o) - y)^2 with respect to the master
parameters (straight-through the BF16 roundings).  The features follow closure_features: q = Pi_neq / rho,
Pi = sqrt(2 q:q), tr q, q:q, tr(q^3).  ``round_bf16=False`` switches the roundings off (for the finite-
difference pins).  lbmpy has no learned closure; the subgrid-model slot it fills is
``turbulence_models.py:25-84`` (Smagorinsky), whose rate formula the network output replaces.
"""
import numpy as np

IN, IN_PAD, H = 10, 16, 32
OFF_W1T, OFF_W2T, OFF_W3, OFF_B1, OFF_B2, OFF_B3, MASTER_SIZE = 0, 512, 1536, 1568, 1600, 1632, 1633


def bf16(a):
    """Round float32 values to BF16 (nearest even) and return them as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(master):
    m = np.asarray(master, dtype=np.float64)
    return dict(w1t=m[OFF_W1T:OFF_W2T].reshape(H, IN_PAD), w2t=m[OFF_W2T:OFF_W3].reshape(H, H),
                w3=m[OFF_W3:OFF_B1], b1=m[OFF_B1:OFF_B2], b2=m[OFF_B2:OFF_B3], b3=m[OFF_B3])


def features(pi6, rho):
    """closure_features for arrays pi6 (n, 6) and rho (n,)."""
    q = np.asarray(pi6, dtype=np.float64) / np.asarray(rho, dtype=np.float64)[:, None]
    xx, yy, zz, xy, xz, yz = q.T
    qq = xx ** 2 + yy ** 2 + zz ** 2 + 2 * (xy ** 2 + xz ** 2 + yz ** 2)
    Q = np.stack([np.stack([xx, xy, xz], -1), np.stack([xy, yy, yz], -1), np.stack([xz, yz, zz], -1)], -2)
    t3 = np.einsum("nab,nbc,nca->n", Q, Q, Q)
    return np.stack([xx, yy, zz, xy, xz, yz, np.sqrt(2 * qq), xx + yy + zz, qq, t3], axis=1)


def softplus(o):
    return np.where(o > 20, o, np.log1p(np.exp(np.minimum(o, 20))))


def sigmoid(o):
    return 1.0 / (1.0 + np.exp(-o))


def forward(master, shift, scale, x, round_bf16=True):
    """Returns o (n,) and the intermediates (a0, a1, h1, h2) for the backward pass."""
    rnd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    p = unpack(master)
    w1t, w2t, w3 = rnd(p["w1t"]), rnd(p["w2t"]), rnd(p["w3"])
    xn = (np.asarray(x, np.float32) - np.asarray(shift, np.float32)) * np.asarray(scale, np.float32)
    a0 = np.zeros((len(x), IN_PAD))
    a0[:, :IN] = rnd(xn)
    h1 = np.tanh(a0 @ w1t.T + p["b1"])
    a1 = rnd(h1)
    h2 = np.tanh(a1 @ w2t.T + p["b2"])
    o = h2 @ w3 + p["b3"]
    return o, (a0, a1, h1, h2, w2t, w3)


def loss_and_grad(master, shift, scale, x, y, round_bf16=True):
    """L = mean((softplus(o) - y)^2) and dL/dmaster (flat, MASTER_SIZE)."""
    n = len(x)
    o, (a0, a1, h1, h2, w2t, w3) = forward(master, shift, scale, x, round_bf16)
    diff = softplus(o) - y
    loss = np.mean(diff ** 2)
    g = 2.0 / n * diff * sigmoid(o)                      # dL/do
    d2 = g[:, None] * w3[None, :] * (1 - h2 ** 2)       # dL/dz2
    d1 = (d2 @ w2t) * (1 - h1 ** 2)                     # dL/dz1
    grad = np.zeros(MASTER_SIZE)
    grad[OFF_W1T:OFF_W2T] = (d1.T @ a0).ravel()
    grad[OFF_W1T:OFF_W2T].reshape(H, IN_PAD)[:, IN:] = 0.0
    grad[OFF_W2T:OFF_W3] = (d2.T @ a1).ravel()
    grad[OFF_W3:OFF_B1] = g @ h2
    grad[OFF_B1:OFF_B2] = d1.sum(0)
    grad[OFF_B2:OFF_B3] = d2.sum(0)
    grad[OFF_B3] = g.sum()
    return loss, grad


def omega(o, nu_scale, omega0, nu_shift=0.0):
    """omega = 1 / max(1/omega0 + 3 (nu_scale softplus(o) - nu_shift), 0.5 + 3e-6)."""
    return 1.0 / np.maximum(1.0 / omega0 + 3.0 * (nu_scale * softplus(o) - nu_shift), 0.500003)
