"""CPU restatement of the learned stress update's network (test infrastructure).

Only tests/ may import this module; it is the checker of the GPU network
(stress-particle-sph_amd/csrc/sps_surrogate.hip), never the thing measured.

The network: 10 -> 128 -> 128 -> 128 -> 4, SiLU on the hidden layers, inputs
padded to 16 and outputs to 16 (the matrix-core tile width), parameters in
the order W1 [16][128], b1, W2 [128][128], b2, W3, b3, W4 [128][16], b4
(weights [in][out]).  Loss: mean over rows and the 4 real outputs of the
squared error.

With bf16=True the GPU's roundings are reproduced: inputs, weights, hidden
activations and the backward pass's output / pre-activation gradients are
rounded to bfloat16 (round to nearest even) where the GPU stores them;
products are summed in single precision (numpy's float32 matmul, in another
order than the matrix cores: results agree to rounding).  With bf16=False the
same function in double precision, which tests pin against central finite
differences.
"""
import numpy as np

NIN, NOUT, IN, H, OUT = 10, 4, 16, 128, 16
SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, OUT)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name]).reshape(-1) for name, _ in SHAPES])


def to_bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, x, bf16=True):
    """Output (n x 4) and the intermediates of a backward pass."""
    dt = np.float32 if bf16 else np.float64
    r = to_bf16 if bf16 else (lambda a: a)
    P = {k: v.astype(dt) for k, v in unpack(np.asarray(theta)).items()}
    n = x.shape[0]
    X = np.zeros((n, IN), dtype=dt)
    X[:, :NIN] = x
    X = r(X)
    Zs, Hs = [], [X]
    Hin = X
    for l in (1, 2, 3):
        Z = Hin @ r(P[f"W{l}"]) + P[f"b{l}"]
        Zs.append(Z)
        Hin = r(silu(Z))
        Hs.append(Hin)
    Y = Hin @ r(P["W4"]) + P["b4"]
    return Y[:, :NOUT], (P, Zs, Hs, Y)


def loss_and_grad(theta, x, t, bf16=True):
    r = to_bf16 if bf16 else (lambda a: a)
    _, (P, Zs, Hs, Y) = forward(theta, x, bf16)
    n = x.shape[0]
    d = Y[:, :NOUT] - t
    loss = float(np.sum(d * d) / (n * NOUT))
    dY = np.zeros_like(Y)
    dY[:, :NOUT] = 2.0 * d / (n * NOUT)
    dOut = r(dY)
    g = {}
    for l in (4, 3, 2, 1):
        Hin = Hs[l - 1]
        g[f"W{l}"] = Hin.T @ dOut
        g[f"b{l}"] = dOut.sum(axis=0)
        if l == 1:
            break
        dH = dOut @ r(P[f"W{l}"]).T
        dOut = r(dH * silu_grad(Zs[l - 2]))
    return loss, pack(g)


def finite_difference(theta, x, t, idx, eps=1e-6):
    """Central differences of the double-precision loss at parameters idx."""
    out = []
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        lp, _ = loss_and_grad(tp, x, t, bf16=False)
        lm, _ = loss_and_grad(tm, x, t, bf16=False)
        out.append((lp - lm) / (2 * eps))
    return np.array(out)
