"""FP64 reference of the planner network (test infrastructure only).

The checker for the library's FP32 MLP kernels (flint_amd/csrc/fa_mlp.hip):
the same architecture (16 -> 64 -> 64 -> 64, ReLU, masked mean square
error) and parameter layout, forward and backward in float64 numpy.  The
planner has no counterpart in the reference (FLINT chooses plans with the
fixed rule mpn_ctx_best_profile, mpn_mul.c:1248-1340); this backward pass is
itself pinned by central finite differences in tests/test_planner_ref.py.
Only tests/ may import this module.
"""
import numpy as np

NI, NH, NO = 16, 64, 64
SHAPES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NO)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SHAPES])


def forward(p, X):
    q = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ q["W1"] + q["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ q["W2"] + q["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ q["W3"] + q["b3"]
    return y, (X, a1, h1, a2, h2)


def loss_grad(p, X, T, M):
    """masked mean square error and its gradient (float64)"""
    q = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    T = np.asarray(T, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    cnt = max(M.sum(), 1.0)
    d = (y - T) * M
    loss = float((d * d).sum() / cnt)
    dy = 2.0 * d / cnt
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    dh2 = (dy @ q["W3"].T) * (a2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ q["W2"].T) * (a1 > 0)
    g["W1"] = X.T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, pack(g)
