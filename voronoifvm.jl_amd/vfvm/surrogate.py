"""Learned Newton initial guess for the nonlinear Poisson family
    eps (u_k - u_l) fluxes, reaction e^u - e^-u - f,
    f a Gaussian random field, Dirichlet values per side.

The network (vfvm_mpnn_*) predicts u from 8 per-node features built here from
the source and the boundary data; the device Newton loop then finishes every
solve to tolerance, so the plain method stays the fallback and the judge of the
prediction.  Sampling, features and the data/eval seeds are defined here so the
training script, the tests and the benchmark use one definition.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _capi
from .grids import linspace, simplexgrid
from .physics import Dirichlet, LinearFlux, Physics, SinhReaction
from .system import Batch, System

NFEAT = 8
GRID_N = 129
EPS = 1e-2
TRAIN_SEEDS = range(0, 10000)
EVAL_SEEDS = range(10000, 11024)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "mpnn_poisson129.npz")


class GRF:
    """Zero-mean Gaussian random field with squared-exponential covariance
    amp^2 exp(-|x - y|^2 / (2 ell^2)) on a tensor grid, sampled exactly through
    the separable factorisation K = Kx (x) Ky (two n x n eigendecompositions)."""

    def __init__(self, n=GRID_N, ell=0.1, amp=5.0):
        x = linspace(0.0, 1.0, n)
        K = np.exp(-((x[:, None] - x[None, :]) ** 2) / (2 * ell * ell))
        w, Q = np.linalg.eigh(K)
        self.L = Q * np.sqrt(np.clip(w, 0.0, None))[None, :]
        self.amp = amp
        self.n = n

    def sample(self, rng):
        Z = rng.standard_normal((self.n, self.n))
        F = self.amp * (self.L @ Z @ self.L.T)  # F[ix, iy]
        return F.ravel(order="F")  # node = ix + iy * n


def draw(seed, grf: GRF):
    """Source (per node) and the four side values (south, east, north, west) of one seed."""
    rng = np.random.default_rng(seed)
    f = grf.sample(rng)
    vals = rng.uniform(-1.0, 1.0, 4)
    return f, vals


class Features:
    """8 per-node input features of the network for sources F (m x n) and side values
    V (m x 4: south, east, north, west): f / 5; asinh(f / 2) (the local balance
    2 sinh u = f); each side value damped by exp(-distance / sqrt(eps / 2)); x; y."""

    def __init__(self, grid, eps=EPS):
        x, y = grid.coord
        delta = math.sqrt(eps / 2.0)
        self.decay = np.stack([np.exp(-d / delta) for d in (y, 1.0 - x, 1.0 - y, x)]).T.astype(np.float32)  # (n, 4)
        self.x, self.y = x.astype(np.float32), y.astype(np.float32)
        self.n = grid.num_nodes

    def __call__(self, F, V):
        F = np.asarray(F, dtype=np.float32).reshape(-1, self.n)
        V = np.asarray(V, dtype=np.float32).reshape(-1, 4)
        m = F.shape[0]
        X = np.empty((m, self.n, NFEAT), dtype=np.float32)
        X[:, :, 0] = F / 5.0
        X[:, :, 1] = np.arcsinh(F / 2.0)
        X[:, :, 2:6] = V[:, None, :] * self.decay[None, :, :]
        X[:, :, 6] = self.x
        X[:, :, 7] = self.y
        return X


def bcs(vals):
    return [Dirichlet(species=1, region=r + 1, value=float(v)) for r, v in enumerate(vals)]


def make_system(grid, f, vals, eps=EPS):
    ph = Physics(flux=LinearFlux(eps), reaction=SinhReaction(), source=np.asfortranarray(f.reshape(1, -1)),
                 bcondition=bcs(vals))
    return System(grid, ph, species=[1])


class Surrogate:
    """vfvm_mpnn handle bound to one grid."""

    def __init__(self, system: System, nlayers=4, seed=0):
        self.system = system
        self.ctx = system.ctx
        h = C.c_void_p()
        _capi.check(_capi.lib().vfvm_mpnn_create(system.handle(), nlayers, 64, seed, C.byref(h)), self.ctx.handle)
        self._h = h
        n = C.c_int64()
        _capi.check(_capi.lib().vfvm_mpnn_params(h, C.byref(n), None))
        self.nparams = n.value
        self.nlayers = nlayers

    def __del__(self):
        if getattr(self, "_h", None) is not None and _capi._LIB is not None:
            _capi._LIB.vfvm_mpnn_destroy(self._h)
            self._h = None

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        _capi.check(_capi.lib().vfvm_mpnn_params(self._h, None, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _capi.check(_capi.lib().vfvm_mpnn_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(self.nparams, np.float32)
        fp = C.POINTER(C.c_float)
        _capi.check(_capi.lib().vfvm_mpnn_loss_grad(self._h, X.shape[0], X.ctypes.data_as(fp), T.ctypes.data_as(fp),
                                                    C.byref(loss), g.ctypes.data_as(fp)), self.ctx.handle)
        return loss.value, g

    def train(self, X, T, epochs, batch=16, lr=1e-3):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float()
        fp = C.POINTER(C.c_float)
        _capi.check(_capi.lib().vfvm_mpnn_train(self._h, X.shape[0], X.ctypes.data_as(fp), T.ctypes.data_as(fp),
                                                epochs, batch, lr, C.byref(loss)), self.ctx.handle)
        return loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], X.shape[1]))
        _capi.check(_capi.lib().vfvm_mpnn_infer(self._h, X.shape[0], X.ctypes.data_as(C.POINTER(C.c_float)),
                                                _capi.dptr(out)), self.ctx.handle)
        return out

    def predict_sides(self, F, V, eps=EPS):
        """Prediction with the features computed on the device from the sources F
        (m x n) and side values V (m x 4); the same features as `Features`."""
        F = np.ascontiguousarray(F, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        out = np.zeros(F.shape)
        _capi.check(_capi.lib().vfvm_mpnn_infer_sides(self._h, F.shape[0], _capi.dptr(F), _capi.dptr(V), eps,
                                                      _capi.dptr(out)), self.ctx.handle)
        return out

    def predict_sides_device(self, F_dev, V_dev, out_dev, nsamples, eps=EPS):
        """predict_sides with the sources, side values and the prediction in device
        memory (vfvm.DeviceArray; out_dev holds nsamples x n float64)."""
        _capi.check(_capi.lib().vfvm_mpnn_infer_sides_device(self._h, int(nsamples), F_dev.ptr, V_dev.ptr, eps,
                                                             out_dev.ptr), self.ctx.handle)
        return out_dev

    def save(self, path=WEIGHTS, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params(), nlayers=self.nlayers, **meta)

    def load(self, path=WEIGHTS):
        d = np.load(path, allow_pickle=False)
        self.set_params(d["params"])
        return d


class PoissonFamily:
    """A batch of `size` systems on one 129 x 129 grid, re-filled seed by seed."""

    def __init__(self, size=256, n=GRID_N):
        X = linspace(0.0, 1.0, n)
        self.grid = simplexgrid(X, X)
        self.grf = GRF(n)
        f0 = np.zeros(self.grid.num_nodes)
        self.systems = [make_system(self.grid, f0, np.zeros(4)) for _ in range(size)]
        self.batch = Batch(self.systems)
        self.size = size
        self.features = Features(self.grid)

    def load_seeds(self, seeds):
        """Fill the batch members with the seeds' data; returns the sources (m x n) and
        side values (m x 4) from which the network's features are computed."""
        seeds = list(seeds)
        assert len(seeds) <= self.size
        F = np.empty((len(seeds), self.grid.num_nodes))
        V = np.empty((len(seeds), 4))
        for m, sd in enumerate(seeds):
            f, vals = draw(sd, self.grf)
            self.batch.update_member(m, source=np.asfortranarray(f.reshape(1, -1)), bcondition=bcs(vals))
            F[m], V[m] = f, vals
        return F, V
