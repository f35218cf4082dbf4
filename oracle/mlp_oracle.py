"""FP64 numpy restatement of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

Network: x -> GELU(x W1 + b1) -> GELU(. W2 + b2) -> . W3 + b3, mean-squared
error over all outputs, Adam (Kingma & Ba 2015).  GELU is the exact
0.5 z (1 + erf(z / sqrt 2)).  With bf16=True the operands the GPU feeds to
its matrix cores (inputs, weights, hidden activations, output and hidden
gradients) are rounded to bfloat16 (round to nearest even) first, so the
GPU result can be compared within FP32-accumulation tolerance.

Pinned by: finite differences of the FP64 loss (gradients), closed forms of
GELU and of a one-step Adam update; the restatements from a GPU trace
(stages_from_trace, grads_from_trace) by the FP64 pipeline itself on inputs
and weights that are exact in bfloat16.
"""
import numpy as np
from scipy.special import erf


def shapes(n_in, h, n_out):
    return [("W1", (n_in, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)),
            ("W3", (h, n_out)), ("b3", (n_out,))]


def unpack(p, n_in, h, n_out):
    out, k = {}, 0
    for name, shp in shapes(n_in, h, n_out):
        n = int(np.prod(shp))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shp)
        k += n
    return out


def pack(d, n_in, h, n_out):
    return np.concatenate([d[name].reshape(-1) for name, _ in shapes(n_in, h, n_out)])


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    v = (((u + r) >> 16) << 16).astype(np.uint32)
    return v.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, x, n_in, h, n_out, bf=False):
    d = unpack(p, n_in, h, n_out)
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    x0 = q(x)
    z1 = x0 @ q(d["W1"]) + d["b1"]
    h1 = q(gelu(z1))
    z2 = h1 @ q(d["W2"]) + d["b2"]
    h2 = q(gelu(z2))
    y = h2 @ q(d["W3"]) + d["b3"]
    return y, (x0, z1, h1, z2, h2)


def loss_and_grads(p, x, t, n_in, h, n_out, bf=False, trace=False):
    """Loss and packed gradients; with trace=True also the intermediates in
    the layout of bd_mlp_trace (z1, h1, z2, h2, y, dy, dyb, dz2, dz2b, dz1,
    dz1b; the *b entries as the matrix cores take them)."""
    d = unpack(p, n_in, h, n_out)
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    y, (x0, z1, h1, z2, h2) = forward(p, x, n_in, h, n_out, bf)
    n = y.size
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / n
    g = {"W3": h2.T @ q(dy), "b3": dy.sum(0)}
    dz2 = (q(dy) @ q(d["W3"]).T) * gelu_grad(z2)
    g["W2"] = h1.T @ q(dz2)
    g["b2"] = dz2.sum(0)
    dz1 = (q(dz2) @ q(d["W2"]).T) * gelu_grad(z1)
    g["W1"] = x0.T @ q(dz1)
    g["b1"] = dz1.sum(0)
    if trace:
        tr = {"z1": z1, "h1": h1, "z2": z2, "h2": h2, "y": y, "dy": dy, "dyb": q(dy),
              "dz2": dz2, "dz2b": q(dz2), "dz1": dz1, "dz1b": q(dz1)}
        return loss, pack(g, n_in, h, n_out), tr
    return loss, pack(g, n_in, h, n_out)


def stages_from_trace(p, x, tr, n_in, h, n_out):
    """Each stage of the GPU's forward and backward recomputed in FP64 from
    the GPU's own inputs to that stage (tr: bd_mlp_trace's arrays): the
    pre-activations z1, z2, y from the BF16 operands, and the FP32
    pre-activation gradients dz2 = (dyb W3^T) gelu'(z2), dz1 = (dz2b W2^T)
    gelu'(z1).  Only the accumulation order can differ from the GPU."""
    d = unpack(p, n_in, h, n_out)
    x0 = bf16(x)
    return {"z1": x0 @ bf16(d["W1"]) + d["b1"],
            "z2": tr["h1"] @ bf16(d["W2"]) + d["b2"],
            "y": tr["h2"] @ bf16(d["W3"]) + d["b3"],
            "dz2": (tr["dyb"] @ bf16(d["W3"]).T) * gelu_grad(tr["z2"]),
            "dz1": (tr["dz2b"] @ bf16(d["W2"]).T) * gelu_grad(tr["z1"])}


def grads_from_trace(x, tr, n_in, h, n_out):
    """Parameter gradients from the GPU's own BF16 operands and FP32
    pre-activation gradients (the products the GPU's matrix cores form), in
    FP64: W3 = h2^T dyb, b3 = sum dy, W2 = h1^T dz2b, b2 = sum dz2,
    W1 = bf16(x)^T dz1b, b1 = sum dz1."""
    g = {"W3": tr["h2"].T @ tr["dyb"], "b3": tr["dy"].sum(0),
         "W2": tr["h1"].T @ tr["dz2b"], "b2": tr["dz2"].sum(0),
         "W1": bf16(x).T @ tr["dz1b"], "b1": tr["dz1"].sum(0)}
    return pack(g, n_in, h, n_out)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
