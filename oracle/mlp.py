"""ORACLE — test infrastructure only.  FP64 numpy restatement of the sizing MLP
(12 -> 64 -> 64 -> 3, ReLU, mean squared error over all outputs) that the HIP
kernels in pathfinder-cpp_amd/csrc/pfh_mlp.hip implement.  Parameter layout:
W1[12][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3] (W[in][out]).

Pinned by central finite differences (tests/test_predictor_oracle.py).
"""
import numpy as np

NI, NH, NO = 12, 64, 3
SHAPES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(params):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def forward(params, x):
    p = unpack(params)
    x = np.asarray(x, np.float64)
    a1 = x @ p["W1"] + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(a2, 0.0)
    out = h2 @ p["W3"] + p["b3"]
    return out, (x, a1, h1, a2, h2)


def loss_and_grad(params, x, target):
    out, (x, a1, h1, a2, h2) = forward(params, x)
    p = unpack(params)
    n = out.shape[0]
    err = out - np.asarray(target, np.float64)
    loss = float((err ** 2).sum() / (n * NO))
    d_out = 2.0 * err / (n * NO)
    g = {"W3": h2.T @ d_out, "b3": d_out.sum(0)}
    d_h2 = (d_out @ p["W3"].T) * (a2 > 0)
    g["W2"] = h1.T @ d_h2
    g["b2"] = d_h2.sum(0)
    d_h1 = (d_h2 @ p["W2"].T) * (a1 > 0)
    g["W1"] = x.T @ d_h1
    g["b1"] = d_h1.sum(0)
    grad = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    return loss, grad


def loss_only(params, x, target):
    out, _ = forward(params, x)
    return float(((out - np.asarray(target, np.float64)) ** 2).sum() / (out.shape[0] * NO))
