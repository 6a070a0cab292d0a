"""Surrogate kernels on the device vs the FP64 CPU oracle MLP (needs a GPU)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import mlp as omlp  # noqa: E402


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("dims,N", [((9, 64, 64, 34), 96), ((9, 512, 512, 512, 512, 1026), 256)])
def test_mlp_forward_backward_vs_fp64(dims, N):
    from sgl_analytic_amd.surrogate import SurrogateNet
    print(f"mlp dims={dims} N={N}", flush=True)
    net = SurrogateNet(dims=dims, seed=3).to("cuda:0")
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(N, dims[0])).astype(np.float32)
    gy = rng.normal(size=(N, dims[-1])).astype(np.float32)
    xt = torch.as_tensor(x, device="cuda:0")
    y = net(xt)
    y.backward(torch.as_tensor(gy, device="cuda:0"))
    torch.cuda.synchronize()
    params = [(W.detach().cpu().double().numpy(), b.detach().cpu().double().numpy())
              for W, b in zip(net.Ws, net.bs)]
    yr, acts, pre = omlp.forward(x.astype(np.float64), params)
    grads = omlp.backward(gy.astype(np.float64), params, acts, pre)
    ey = _rel(y.detach().cpu().numpy(), yr)
    print("forward rel err", ey)
    assert ey < 2e-2          # BF16 weights/activations, FP32 accumulation
    for i, (W, b) in enumerate(zip(net.Ws, net.bs)):
        eW = _rel(W.grad.cpu().numpy(), grads[i][0])
        eb = _rel(b.grad.cpu().numpy(), grads[i][1])
        print(f"layer {i}: dW rel err {eW:.3e} db rel err {eb:.3e}")
        assert eW < 5e-2 and eb < 5e-2


def test_lnphi_interp_is_cubic_exact():
    """Cubic Lagrange interpolation reproduces a cubic in the node index
    exactly (ln k is uniform in the node index)."""
    import ctypes
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import Shape, _ptr, _stream
    sh = Shape(n_k=1024)
    n_c = 64
    u = np.arange(n_c, dtype=np.float64)
    f = lambda t: 0.3 - 0.2 * t + 0.01 * t * t - 1e-4 * t ** 3
    y = np.zeros((2, 2 * n_c), np.float32)
    y[0, 0::2] = f(u); y[0, 1::2] = -f(u)
    y[1, 0::2] = 2 * f(u); y[1, 1::2] = 0.5
    yt = torch.as_tensor(y, device="cuda:0")
    lam = torch.zeros((2, sh.n_k, 2), dtype=torch.float64, device="cuda:0")
    cs = sh.c()
    rc = _lib.load().sgl_lnphi_interp(_ptr(yt), 2 * n_c, n_c, 2, ctypes.byref(cs), _ptr(lam), _stream())
    assert rc == 0
    torch.cuda.synchronize()
    t = np.arange(sh.n_k) * (n_c - 1) / (sh.n_k - 1)
    L = lam.cpu().numpy()
    yf = y.astype(np.float64)
    ref = f(t)
    # interpolating the FP32-rounded node values: exact up to FP32 rounding
    assert np.max(np.abs(L[0, :, 0] - ref)) < 1e-5
    assert np.max(np.abs(L[0, :, 1] + ref)) < 1e-5
    assert np.max(np.abs(L[1, :, 1] - 0.5)) < 1e-6


def _trained():
    import os
    from sgl_analytic_amd import surrogate as SU
    path = SU.default_weights()
    if not os.path.exists(path):
        pytest.skip("no trained surrogate weights in sgl_analytic_amd/data")
    return SU, SU.load(path)


def test_trained_network_gain_and_gates():
    """On cosmologies the networks never saw, the trained networks must cut
    the cheap solve's L1 error against the batched solver, and every PDF the
    gates accept must be within 1e-3 L1 of the solver."""
    from sgl_analytic_amd import solver as S
    SU, nets = _trained()
    B = 1024
    params = SU.lhs_params(B, seed=4242)
    n_z = S.slice_count(params[:, 9])
    shape = S.Shape(n_z_max=82)
    rec = S.cosmo_records(params, n_z)
    print(f"surrogate gain B={B} seed=4242", flush=True)
    x = torch.as_tensor(SU.features(params), dtype=torch.float32, device="cuda:0")
    inf = SU.Inference(nets, shape, B, n_z_max=82)
    inf.set_records(rec)
    p, d, acc, y, _ = inf.run(x)
    torch.cuda.synchronize()
    p_net = p.cpu().numpy().astype(np.float64)
    acc = acc.cpu().numpy()
    p0, _, _, _, _ = inf.run(x, use_network=False)
    torch.cuda.synchronize()
    p_cheap = p0.cpu().numpy().astype(np.float64)
    del inf
    sv = S.Solver(shape, B, path="binned")
    pr, _ = sv.run(rec)
    torch.cuda.synchronize()
    pr = pr.float().cpu().numpy().astype(np.float64)
    xo = shape.out_grid()
    l1 = np.trapezoid(np.abs(p_net - pr), xo, axis=1)
    l1c = np.trapezoid(np.abs(p_cheap - pr), xo, axis=1)
    print("median L1 with networks", np.median(l1), "without", np.median(l1c),
          "accept", acc.mean(), "accepted max", l1[acc].max() if acc.any() else None)
    assert np.median(l1) < 0.25 * np.median(l1c)
    assert np.quantile(l1, 0.9) < 0.5 * np.quantile(l1c, 0.9)
    assert acc.mean() > 0.5
    assert np.all(l1[acc] <= 1e-3)
