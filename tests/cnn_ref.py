"""NumPy reference of the learned preconditioner network (test infrastructure).

Same architecture, parameter layout and padding rules as csrc/pv_cnn.hip:
6 layers of 3x3 convolutions (3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1), ReLU
after layers 1..5, NHWC tensors, zero padding in y, periodic or zero padding
in x.  With bf16=True the tensors the GPU stores in BF16 (activations of
layers 1..5, the 32 -> 32 weights, the pre-activation gradients) are rounded
to BF16 at the same points, everything else is float64.  The backward pass is
pinned by finite differences in tests/test_cnn_ref.py.
"""
import numpy as np

NPARAM = 38177
C = 32


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    """Parameter vector -> six (W[tap][co][ci], b[co]) pairs."""
    p = np.asarray(p, dtype=np.float64)
    layers = [(p[0:864].reshape(32, 9, 3).transpose(1, 0, 2), p[864:896])]
    for k in range(4):
        o = 896 + k * 9248
        layers.append((p[o:o + 9216].reshape(9, 32, 32), p[o + 9216:o + 9248]))
    layers.append((p[37888:38176].reshape(9, 1, 32), p[38176:38177]))
    return layers


def pack_grad(gl):
    """Six (dW[tap][co][ci], db) pairs -> gradient vector in parameter layout."""
    g = np.zeros(NPARAM)
    g[0:864] = gl[0][0].transpose(1, 0, 2).reshape(-1)
    g[864:896] = gl[0][1]
    for k in range(4):
        o = 896 + k * 9248
        g[o:o + 9216] = gl[k + 1][0].reshape(-1)
        g[o + 9216:o + 9248] = gl[k + 1][1]
    g[37888:38176] = gl[5][0].reshape(-1)
    g[38176] = gl[5][1][0]
    return g


def pad(x, px):
    x = np.pad(x, ((0, 0), (1, 1), (0, 0), (0, 0)))
    if px:
        return np.concatenate([x[:, :, -1:], x, x[:, :, :1]], axis=2)
    return np.pad(x, ((0, 0), (0, 0), (1, 1), (0, 0)))


def conv(x, W, b, px):
    """out[b, y, x, co] = b[co] + sum_tap sum_ci W[tap][co][ci] x[b, y+ky-1, x+kx-1, ci]."""
    B, H, Wd, _ = x.shape
    xp = pad(x, px)
    out = np.zeros((B, H, Wd, W.shape[1])) + b
    for t in range(9):
        ky, kx = divmod(t, 3)
        out += xp[:, ky:ky + H, kx:kx + Wd, :] @ W[t].T
    return out


def conv_back(x, W, gout, px):
    """Gradients of conv w.r.t. W, b and x for the upstream gradient gout."""
    B, H, Wd, _ = x.shape
    xp = pad(x, px)
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    for t in range(9):
        ky, kx = divmod(t, 3)
        dW[t] = np.einsum("bhwo,bhwi->oi", gout, xp[:, ky:ky + H, kx:kx + Wd, :])
        dxp[:, ky:ky + H, kx:kx + Wd, :] += gout @ W[t]
    dx = dxp[:, 1:H + 1]
    if px:
        dx = dx[:, :, 1:Wd + 1] + 0.0
        dx[:, :, -1] += dxp[:, 1:H + 1, 0]
        dx[:, :, 0] += dxp[:, 1:H + 1, Wd + 1]
    else:
        dx = dx[:, :, 1:Wd + 1]
    return dW, gout.sum(axis=(0, 1, 2)), dx


def forward(p, x, px, use_bf16=False):
    """x [B][H][W][3] -> out [B][H][W] and the activations for backward."""
    L = unpack(p)
    rnd = bf16 if use_bf16 else (lambda a: a)
    acts = [np.asarray(x, dtype=np.float64)]
    a = np.maximum(conv(acts[0], L[0][0], L[0][1], px), 0.0)
    acts.append(rnd(a))
    for k in range(1, 5):
        a = np.maximum(conv(acts[-1], rnd(L[k][0]), L[k][1], px), 0.0)
        acts.append(rnd(a))
    out = conv(acts[-1], L[5][0], L[5][1], px)[..., 0]
    return out, acts


def loss_and_gout(out, tgt, border):
    B, H, W = out.shape
    m = np.zeros_like(out)
    m[:, border:H - border, border:W - border] = 1.0
    n = max(m.sum(), 1.0)
    d = (out - tgt) * m
    return float((d * d).sum() / n), 2.0 * d / n


def backward(p, acts, gout, px, use_bf16=False):
    L = unpack(p)
    rnd = bf16 if use_bf16 else (lambda a: a)
    grads = [None] * 6
    dW, db, da = conv_back(acts[5], L[5][0], gout[..., None], px)
    grads[5] = (dW, db)
    gz = rnd(da * (acts[5] > 0))
    for k in range(4, 0, -1):
        dW, db, da = conv_back(acts[k], rnd(L[k][0]), gz, px)
        grads[k] = (dW, db)
        gz = rnd(da * (acts[k] > 0))
    dW, db, _ = conv_back(acts[0], L[0][0], gz, px)
    grads[0] = (dW, db)
    return pack_grad(grads)


def init_params(seed=1):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[0:864] = rng.standard_normal(864) * np.sqrt(2 / 27)
    for k in range(4):
        o = 896 + k * 9248
        p[o:o + 9216] = rng.standard_normal(9216) * np.sqrt(2 / 288)
        p[o + 9216:o + 9248] = 0.01 * rng.standard_normal(32)
    p[864:896] = 0.01 * rng.standard_normal(32)
    p[37888:38176] = 0.1 * rng.standard_normal(288) * np.sqrt(2 / 288)
    return p
