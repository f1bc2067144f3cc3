"""TEST INFRASTRUCTURE -- numpy restatement of the rank-prediction MLP
(tensor_train_amd/csrc/tta_mlp.hip) for gradient checks: n_in -> H x L
(ReLU) -> 1, mean-squared error, parameters in the library's flat layout
(per layer W (in x out, row-major) then b (out)).  The MLP has no
counterpart in the reference; this file pins the device forward/backward
against plain float64 arithmetic."""
import numpy as np


def unpack(p, dims):
    Ws, bs, off = [], [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        Ws.append(p[off:off + a * b].reshape(a, b).astype(np.float64))
        off += a * b
        bs.append(p[off:off + b].astype(np.float64))
        off += b
    return Ws, bs


def forward(p, dims, X):
    Ws, bs = unpack(p, dims)
    acts = [X.astype(np.float64)]
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = acts[-1] @ W + b
        acts.append(np.maximum(z, 0.0) if i < len(Ws) - 1 else z)
    return acts


def loss_and_grad(p, dims, X, y):
    Ws, _ = unpack(p, dims)
    acts = forward(p, dims, X)
    n = X.shape[0]
    out = acts[-1][:, 0]
    err = out - y.astype(np.float64)
    loss = np.mean(err ** 2)
    d = (2.0 * err / n)[:, None]
    grads = [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gW = acts[l].T @ d
        gb = d.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            d = (d @ Ws[l].T) * (acts[l] > 0)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def round_bf16(x):
    """Nearest bfloat16 value of each float32 entry (ties to even), returned
    as float32: the upper 16 bits of the IEEE single after adding 0x7fff plus
    the lowest kept bit; infinities and NaNs pass through."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    finite = (b & 0x7F800000) != 0x7F800000
    r = np.where(finite, b + 0x7FFF + ((b >> 16) & 1), b) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def forward_bf16(p, dims, X):
    """Inference output with weights, biases and layer inputs rounded to
    bfloat16 and products accumulated exactly (float64)."""
    Ws, bs = unpack(round_bf16(np.asarray(p, np.float32)), dims)
    a = X
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = round_bf16(np.asarray(a, np.float32)).astype(np.float64) @ W + b
        a = np.maximum(z, 0.0) if i < len(Ws) - 1 else z
    return a[:, 0]
