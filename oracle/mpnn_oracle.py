"""CPU oracle of the message-passing network that predicts the Newton initial
iterate: TEST INFRASTRUCTURE (the checker of the GPU training and inference
kernels).  Only tests/ and bench.py's cpu_baseline leg import it.

The reference has no learned component; this network is defined by this
project (the "learned Newton initial guess" of the hot-path specification),
so the oracle restates the project's own definition in plain numpy:

    X  : n x 16 network input: 8 node features, then context channels 8..15 =
         features 1 and 0 smoothed by K = 16, 64, 160, 324 steps of
         y <- (y + W y) / 2, W y_i = sum_j w_ij y_j (snapshots in that order)
    h0 = relu(bf(X) @ bf(We) + be)                                  (n x 64)
    for l in 0..L-1:
        a_l  = Agg(h_l),  Agg(h)_i = sum_j w_ij h_j - h_i,  w_ij = s_ij / sum_k s_ik
        c_l  = [bf(h_l), bf(a_l)]                                   (n x 128)
        r_l  = relu(c_l @ bf(W_l) + b_l)
        h_{l+1} = h_l + r_l
    u  = h_L @ wd + bd
    loss = mean over rows of (u - target)^2

s_ij is the Voronoi edge factor sigma/h of the edge (i, j).  bf() rounds to
bfloat16 (round to nearest even), exactly where the GPU kernels feed the
matrix cores; everything else is evaluated in float64 here (the GPU
accumulates in float32).  With emulate_bf16=False the same code is the exact
float64 function, which the finite-difference test pins.

Parameter vector layout (float32 on the device), in order:
    We (16 x 64), be (64), [W_l (128 x 64), b_l (64)] * L, wd (64), bd (1)
"""
from __future__ import annotations

import numpy as np

NFEAT = 8
KPAD = 16
WIDTH = 64


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32/float64 values to bfloat16 precision (round to nearest even)."""
    x = np.asarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    r = ((u + rounding) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(x)
    if nan.any():
        out = np.where(nan, x, out)
    return out.astype(np.float64)


def nparams(nlayers=4, width=WIDTH):
    return KPAD * width + width + nlayers * (2 * width * width + width) + width + 1


def unpack(theta, nlayers=4, width=WIDTH):
    theta = np.asarray(theta, dtype=np.float64)
    p = 0

    def take(shape):
        nonlocal p
        n = int(np.prod(shape))
        v = theta[p:p + n].reshape(shape)
        p += n
        return v

    We = take((KPAD, width))
    be = take((width,))
    layers = []
    for _ in range(nlayers):
        W = take((2 * width, width))
        b = take((width,))
        layers.append((W, b))
    wd = take((width,))
    bd = take((1,))
    assert p == theta.size
    return We, be, layers, wd, bd


def init_params(seed, nlayers=4, width=WIDTH):
    """He-uniform initialisation (the library uses the same rule, its own RNG)."""
    rng = np.random.default_rng(seed)
    parts = []
    parts.append(rng.uniform(-1, 1, (KPAD, width)) * np.sqrt(6.0 / KPAD))
    parts.append(np.zeros(width))
    for _ in range(nlayers):
        parts.append(rng.uniform(-1, 1, (2 * width, width)) * np.sqrt(6.0 / (2 * width)) * 0.5)
        parts.append(np.zeros(width))
    parts.append(rng.uniform(-1, 1, width) * np.sqrt(3.0 / width))
    parts.append(np.zeros(1))
    return np.concatenate([p.ravel() for p in parts])


def graph_weights(adj_ptr, adj_col, adj_fac):
    """Row-normalised edge weights w_ij = s_ij / sum_k s_ik (self entries dropped)."""
    n = adj_ptr.size - 1
    rows, cols, w = [], [], []
    for i in range(n):
        sl = slice(adj_ptr[i], adj_ptr[i + 1])
        c = adj_col[sl]
        f = adj_fac[sl]
        keep = c != i
        s = f[keep].sum()
        rows.append(np.full(keep.sum(), i))
        cols.append(c[keep])
        w.append(f[keep] / s if s > 0 else f[keep] * 0)
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(w)


SMOOTH_STEPS = (16, 64, 160, 324)
SMOOTH_SRC = (1, 0)


def expand(X, graph):
    """(nsamples, n, 8) features -> (nsamples, n, 16) network input with the smoothed
    context channels (float64 here; the device smooths in float32)."""
    import scipy.sparse as sp
    rows, cols, w = graph
    ns, n, _ = X.shape
    Xp = np.zeros((ns, n, KPAD))
    Xp[:, :, :NFEAT] = X
    Wm = sp.csr_matrix((w, (rows, cols)), shape=(n, n))
    for c, src in enumerate(SMOOTH_SRC):
        y = Xp[:, :, src].T.copy()
        snap = 0
        for step in range(1, SMOOTH_STEPS[-1] + 1):
            y = 0.5 * (y + Wm @ y)
            if step == SMOOTH_STEPS[snap]:
                Xp[:, :, NFEAT + 4 * c + snap] = y.T
                snap += 1
    return Xp


def _agg(h, rows, cols, w, n):
    out = np.zeros_like(h)
    np.add.at(out, rows, w[:, None] * h[cols])
    return out - h


def _aggT(d, rows, cols, w, n):
    out = np.zeros_like(d)
    np.add.at(out, cols, w[:, None] * d[rows])
    return out - d


def forward(theta, X, graph, nlayers=4, emulate_bf16=True, keep=False):
    """X: (nsamples, n, NFEAT) features.  Returns u (nsamples, n) [and the cache]."""
    bf = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    We, be, layers, wd, bd = unpack(theta, nlayers)
    rows, cols, w = graph
    ns, n, _ = X.shape
    Xp = expand(X, graph)
    us, caches = [], []
    for s in range(ns):
        x = bf(Xp[s])
        z0 = x @ bf(We) + be
        h = np.maximum(z0, 0.0)
        cache = dict(x=x, h0=h.copy(), c=[], r=[])
        for (W, b) in layers:
            a = _agg(h, rows, cols, w, n)
            c = np.concatenate([bf(h), bf(a)], axis=1)
            r = np.maximum(c @ bf(W) + b, 0.0)
            cache["c"].append(c)
            cache["r"].append(r)
            h = h + r
        cache["hL"] = h
        u = h @ wd + bd[0]
        us.append(u)
        caches.append(cache)
    u = np.array(us)
    return (u, caches) if keep else u


def loss_and_grad(theta, X, targets, graph, nlayers=4, emulate_bf16=True):
    """Mean squared error over all rows and its gradient with respect to theta.
    The backward pass differentiates the float64 function through the rounded
    operands (bf() is treated as the identity, as the GPU kernels do)."""
    bf = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    We, be, layers, wd, bd = unpack(theta, nlayers)
    rows, cols, w = graph
    u, caches = forward(theta, X, graph, nlayers, emulate_bf16, keep=True)
    ns, n = u.shape
    diff = u - targets
    loss = float((diff ** 2).mean())
    du_all = 2.0 * diff / diff.size
    gWe, gbe = np.zeros_like(We), np.zeros_like(be)
    gL = [(np.zeros_like(W), np.zeros_like(b)) for (W, b) in layers]
    gwd, gbd = np.zeros_like(wd), np.zeros(1)
    for s in range(ns):
        cache = caches[s]
        du = du_all[s]
        gwd += cache["hL"].T @ du
        gbd += du.sum()
        dh = np.outer(du, wd)
        for l in range(nlayers - 1, -1, -1):
            W, b = layers[l]
            c, r = cache["c"][l], cache["r"][l]
            dz = dh * (r > 0)
            gL[l][0][...] += bf(c).T @ bf(dz) if emulate_bf16 else c.T @ dz
            gL[l][1][...] += dz.sum(axis=0)
            dc = (bf(dz) if emulate_bf16 else dz) @ bf(W).T
            dh = dh + dc[:, :WIDTH] + _aggT(dc[:, WIDTH:], rows, cols, w, n)
        dz0 = dh * (cache["h0"] > 0)
        gWe += cache["x"].T @ (bf(dz0) if emulate_bf16 else dz0)
        gbe += dz0.sum(axis=0)
    parts = [gWe, gbe]
    for (gW, gb) in gL:
        parts += [gW, gb]
    parts += [gwd, gbd]
    return loss, np.concatenate([p.ravel() for p in parts])
