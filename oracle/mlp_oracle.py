"""CPU oracle of the sigma(B) surrogate MLP (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU leg may import this
module, and only as the checker.  It restates, in float64 numpy, what the MFMA
kernels of elasticscattering_amd/csrc/es_mlp.hip compute:

    y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3,   6 -> 128 -> 128 -> 2,
    loss = mean_rows sum_o w_o (y_o - t_o)^2,

its exact gradient (``emulate_bf16=False``), the same gradient with every
matrix-product operand rounded where the kernels round it
(``emulate_bf16=True``: the weights, dy, dz2 and dz1 to bfloat16; x, h1, h2
to a bfloat16 pair hi + lo in the forward products and to hi in the weight
gradients — FP32 accumulation is then the only remaining difference), and one
Adam step
(Kingma & Ba 2015, bias-corrected).  The reference has no counterpart; the
exact gradient is pinned by central finite differences in the tests.

Parameter vector layout (one member, 20752 floats, padded):
W1[128][16] | b1[128] | W2[128][128] | b2[128] | W3[16][128] | b3[16].
"""
import numpy as np

N_IN, N_INP, N_H, N_OUTP = 6, 16, 128, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_H * N_INP
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_OUTP * N_H
NPARAM = OFF_B3 + N_OUTP


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def split_bf16(a):
    """hi + lo with hi = bf16(a), lo = bf16(a - hi), a taken to float32 first."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    hi = bf16(a32)
    lo = bf16((a32.astype(np.float64) - hi).astype(np.float32))
    return hi + lo


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_H, N_INP)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_H, N_H)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(N_OUTP, N_H)
    b3 = p[OFF_B3:NPARAM]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def _pad_x(x):
    x = np.asarray(x, dtype=np.float64)
    xp = np.zeros((x.shape[0], N_INP))
    xp[:, :N_IN] = x
    return xp


def forward(p, x, emulate_bf16=True):
    """y [n][2] and the activations (x, h1, h2 as the weight gradients use them) of one member."""
    ident = (lambda a: np.asarray(a, dtype=np.float64))
    r = bf16 if emulate_bf16 else ident
    s = split_bf16 if emulate_bf16 else ident
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xp = _pad_x(x)
    h1 = np.tanh(s(xp) @ r(W1).T + b1)
    h2 = np.tanh(s(h1) @ r(W2).T + b2)
    y = s(h2) @ r(W3).T + b3
    return y[:, :2], (r(xp), h1, r(h1), h2, r(h2))


def loss_and_grad(p, x, t, out_weight=(1.0, 1.0), emulate_bf16=True):
    """(loss, gradient [NPARAM]) of one member over the rows x [n][6], t [n][2]."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = x.shape[0]
    ow = np.asarray(out_weight, dtype=np.float64)
    y, (xb, h1, h1b, h2, h2b) = forward(p, x, emulate_bf16)
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(ow * e * e) / n)
    dy = np.zeros((n, N_OUTP))
    dy[:, :2] = 2.0 * ow * e / n
    dyb = r(dy)
    gW3 = dyb.T @ h2b
    gb3 = dy.sum(axis=0)
    dz2 = (dyb @ r(W3)) * (1.0 - h2 * h2)
    dz2b = r(dz2)
    gW2 = dz2b.T @ h1b
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2b @ r(W2)) * (1.0 - h1 * h1)
    dz1b = r(dz1)
    gW1 = dz1b.T @ xb
    gb1 = dz1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam_step(p, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam step; returns (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def finite_difference(p, x, t, idx, h=1e-6, out_weight=(1.0, 1.0)):
    """Central differences of the exact loss at parameter indices idx."""
    out = []
    for i in idx:
        pp = np.array(p, dtype=np.float64)
        pm = np.array(p, dtype=np.float64)
        pp[i] += h
        pm[i] -= h
        lp, _ = loss_and_grad(pp, x, t, out_weight, emulate_bf16=False)
        lm, _ = loss_and_grad(pm, x, t, out_weight, emulate_bf16=False)
        out.append((lp - lm) / (2 * h))
    return np.array(out)
