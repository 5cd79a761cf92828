"""The initial-guess network's GPU kernels against the CPU oracle (bfloat16
rounding emulated at the same points), and its training behaviour."""
import os

import numpy as np
import pytest

import mpnn_oracle as M
import vfvm
from vfvm import surrogate as S

pytestmark = pytest.mark.gpu


def _setup(n=9, nsamples=3, seed=0, nlayers=4):
    X = vfvm.linspace(0.0, 1.0, n)
    grid = vfvm.simplexgrid(X, X)
    sys_ = S.make_system(grid, np.zeros(grid.num_nodes), np.zeros(4))
    net = S.Surrogate(sys_, nlayers=nlayers, seed=seed)
    import ctypes as C
    from vfvm import _capi
    h = sys_.handle()
    # the oracle's graph from the same node graph the library uses
    import oracle as O
    o = O.OracleSystem(grid, 1, O.physics("linear", D0=1.0), {}, O.EDGEWISE)
    en, ef = o.edges()
    n_ = grid.num_nodes
    adj = [[] for _ in range(n_)]
    for (a, b), f in zip(en, ef):
        if f != 0.0:
            adj[a].append((b, f))
            adj[b].append((a, f))
    ptr, col, fac = [0], [], []
    for i in range(n_):
        for j, f in sorted(adj[i]):
            col.append(j)
            fac.append(f)
        ptr.append(len(col))
    graph = M.graph_weights(np.array(ptr), np.array(col), np.array(fac))
    rng = np.random.default_rng(seed)
    feats = rng.normal(size=(nsamples, n_, M.NFEAT)).astype(np.float32)
    targets = rng.normal(size=(nsamples, n_)).astype(np.float32)
    return net, graph, feats, targets


def test_forward_matches_oracle():
    print("case mpnn forward", flush=True)
    net, graph, X, T = _setup()
    theta = net.params().astype(np.float64)
    u_gpu = net.predict(X)
    u_cpu = M.forward(theta, X.astype(np.float64), graph, nlayers=4, emulate_bf16=True)
    scale = np.abs(u_cpu).max()
    err = np.abs(u_gpu - u_cpu)
    print("forward max rel err", err.max() / scale, "rms", np.sqrt((err ** 2).mean()) / scale, flush=True)
    # float32 accumulation vs float64; an activation may round to the neighbouring bfloat16
    assert err.max() <= 2e-2 * scale
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * scale


def test_forward_matches_oracle_many_tiles():
    """169,000 rows: more 64-row tiles than inference blocks (blocks loop over tiles),
    a partial last tile, and 80 smoothing columns (two column blocks)."""
    print("case mpnn forward, 65x65 grid, 40 samples", flush=True)
    net, graph, X, T = _setup(n=65, nsamples=40, seed=3)
    theta = net.params().astype(np.float64)
    u_gpu = net.predict(X)
    u_cpu = M.forward(theta, X.astype(np.float64), graph, nlayers=4, emulate_bf16=True)
    scale = np.abs(u_cpu).max()
    err = np.abs(u_gpu - u_cpu)
    print("forward max rel err", err.max() / scale, "rms", np.sqrt((err ** 2).mean()) / scale, flush=True)
    assert err.max() <= 2e-2 * scale
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * scale


def test_gradient_matches_oracle():
    print("case mpnn gradient", flush=True)
    net, graph, X, T = _setup(seed=1)
    theta = net.params().astype(np.float64)
    loss_g, grad_g = net.loss_grad(X, T)
    loss_c, grad_c = M.loss_and_grad(theta, X.astype(np.float64), T.astype(np.float64), graph, nlayers=4,
                                     emulate_bf16=True)
    rel = np.linalg.norm(grad_g - grad_c) / np.linalg.norm(grad_c)
    cos = grad_g @ grad_c / (np.linalg.norm(grad_g) * np.linalg.norm(grad_c))
    print("loss", loss_g, loss_c, "grad rel err", rel, "cos", cos, flush=True)
    assert abs(loss_g - loss_c) <= 1e-3 * abs(loss_c)
    assert rel <= 2e-2 and cos >= 0.9995
    # blockwise: every parameter group agrees
    We, be, layers, wd, bd = M.unpack(grad_c)
    Wg, bg, lg, wdg, bdg = M.unpack(grad_g.astype(np.float64))
    for a, b in [(We, Wg), (be, bg), (wd, wdg)] + [(x[0], y[0]) for x, y in zip(layers, lg)]:
        assert np.linalg.norm(a - b) <= 3e-2 * max(np.linalg.norm(a), 1e-12)


def test_training_reduces_loss():
    print("case mpnn training", flush=True)
    net, graph, X, T = _setup(n=17, nsamples=8, seed=2)
    # a learnable target: a smooth function of the features
    T = (0.5 * np.tanh(X[:, :, 0]) + 0.25 * X[:, :, 1] * X[:, :, 2]).astype(np.float32)
    l0, _ = net.loss_grad(X, T)
    l1 = net.train(X, T, epochs=150, batch=4, lr=3e-3)
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.05 * l0


@pytest.mark.skipif(not os.path.exists(S.WEIGHTS), reason="trained weights not present")
def test_trained_surrogate_saves_newton_iterations():
    """What the trained network gains: fewer Newton iterations to ||F|| <= 1e-10 on
    held-out seeds than from a zero initial value, with the same converged answer."""
    print("case surrogate gain", flush=True)
    fam = S.PoissonFamily(size=16)
    net = S.Surrogate(fam.systems[0])
    net.load()
    seeds = list(S.EVAL_SEEDS)[:16]
    raw = fam.load_seeds(seeds)
    feats = fam.features(*raw)
    ctrl = vfvm.SolverControl(precision="fp32", reltol_linear=1e-6, abstol_linear=0.0, maxiters_linear=2000)
    U0 = fam.batch.solve(0.0, ctrl)
    it0 = [h.niter for h in fam.batch.histories]
    guess = net.predict(feats)
    Ug = fam.batch.solve([g.reshape(1, -1) for g in guess], ctrl)
    itg = [h.niter for h in fam.batch.histories]
    # the device feature path gives the same prediction (bfloat16 input rounding aside)
    guess_dev = net.predict_sides(*raw)
    dev_diff = np.abs(guess_dev - guess).max()
    print("device-feature prediction max diff", dev_diff, flush=True)
    assert dev_diff <= 2e-2 * np.abs(guess).max()
    # inputs and prediction in device memory: the same numbers as the host-buffer call
    ctx = fam.systems[0].ctx
    F_dev = vfvm.DeviceArray.from_numpy(ctx, raw[0])
    V_dev = vfvm.DeviceArray.from_numpy(ctx, raw[1])
    out_dev = vfvm.DeviceArray(ctx, raw[0].size)
    net.predict_sides_device(F_dev, V_dev, out_dev, len(seeds))
    assert np.array_equal(out_dev.numpy().reshape(guess_dev.shape), guess_dev)
    err = max(np.abs(g - vfvm.dofs(u)).max() for g, u in zip(guess, U0))
    diff = max(np.abs(vfvm.dofs(a) - vfvm.dofs(b)).max() for a, b in zip(U0, Ug))
    print("newton iterations from zero", sum(it0), "from surrogate", sum(itg), "surrogate max err", err, flush=True)
    assert max(h.residual_norm for h in fam.batch.histories) <= 1e-10
    assert diff <= 1e-9
    assert sum(itg) < sum(it0)
