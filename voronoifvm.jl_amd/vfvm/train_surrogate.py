"""Trains the initial-guess network on plain GPU Newton solutions.

    python -m vfvm.train_surrogate [--samples 10000] [--epochs 24]

Targets are converged solutions of the nonlinear Poisson family (129 x 129 grid,
eps = 1e-2, GRF source, side values) for the training seeds, computed by the
batched FP64 device Newton loop; the network is trained by the library's own
forward/backward kernels (Adam, mini-batches of 16).  Writes the parameters to
vfvm/data/mpnn_poisson129.npz.
"""
import argparse
import sys
import time

import numpy as np

from . import surrogate as S
from .control import SolverControl


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--out", default=S.WEIGHTS, help="where to write the trained parameters")
    args = ap.parse_args()
    t0 = time.time()
    fam = S.PoissonFamily(size=256)
    seeds = list(S.TRAIN_SEEDS)[:args.samples]
    X = np.empty((len(seeds), fam.grid.num_nodes, S.NFEAT), np.float32)
    T = np.empty((len(seeds), fam.grid.num_nodes), np.float32)
    ctrl = SolverControl(precision="fp64", reltol_linear=1e-8, abstol_linear=0.0, maxiters_linear=3000)
    for b0 in range(0, len(seeds), 256):
        blk = seeds[b0:b0 + 256]
        X[b0:b0 + len(blk)] = fam.features(*fam.load_seeds(blk))
        U = fam.batch.solve(0.0, ctrl)
        worst = max(h.residual_norm for h in fam.batch.histories[:len(blk)])
        assert worst <= 1e-10, worst
        for k in range(len(blk)):
            T[b0 + k] = U[k].ravel()
        print(f"data {b0 + len(blk)}/{len(seeds)} ({time.time() - t0:.1f}s)", flush=True)
    net = S.Surrogate(fam.systems[0], nlayers=4, seed=0)
    schedule = [(args.epochs // 2, 2e-3), (args.epochs // 4, 7e-4), (args.epochs - args.epochs // 2 - args.epochs // 4, 2e-4)]
    loss = None
    for ep, lr in schedule:
        for e in range(ep):
            loss = net.train(X, T, epochs=1, batch=args.batch, lr=lr)
            print(f"epoch lr={lr:g} loss={loss:.3e} ({time.time() - t0:.1f}s)", flush=True)
    net.save(args.out, samples=len(seeds), epochs=args.epochs, final_loss=loss)
    print(f"saved {args.out} loss={loss:.3e} total {time.time() - t0:.1f}s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
