"""Two-point term of the batched path on the k sub-grid (needs a GPU).

sgl_psi2_interp is four-point Lagrange in the node index: it must pass the
sub-grid values through bit-exactly and reproduce any cubic in the node
index (the rule is exact for cubics), including the one-sided first and
last intervals.  The batched solver's Psi_2 on every node is the
interpolated sub-grid result."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _interp(y, n_kc, n_k, B):
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import Shape, _ptr, _stream
    L = _lib.load()
    cs = Shape(n_k=n_k).c()
    assert L.sgl_psi2_nodes(ctypes.byref(cs)) == n_kc
    yc = torch.as_tensor(y, dtype=torch.float64).cuda().contiguous()
    out = torch.full((B, n_k, 2), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(L.sgl_psi2_interp(_ptr(yc), B, ctypes.byref(cs), _ptr(out), _stream()),
               "sgl_psi2_interp")
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("n_k", [8192, 1024, 513, 25])
def test_interp_passes_nodes_and_cubics(n_k):
    from sgl_analytic_amd._lib import PSI2_STRIDE as S
    n_kc = (n_k - 1 + S - 1) // S + 1
    B = 3
    print(f"psi2 interp n_k={n_k} n_kc={n_kc}", flush=True)
    rng = np.random.default_rng(7)
    y = rng.standard_normal((B, n_kc, 2))
    out = _interp(y, n_kc, n_k, B)
    nodes = np.arange(0, n_k, S)
    assert np.array_equal(out[:, nodes, :], y[:, : len(nodes), :])
    # cubic in the node index n: reproduced to round-off everywhere
    c = rng.standard_normal((B, 2, 4))
    jc = np.arange(n_kc) * S / (n_k - 1.0)
    n = np.arange(n_k) / (n_k - 1.0)
    cub = lambda x: sum(c[:, :, p, None] * x[None, None, :] ** p for p in range(4))
    yc = np.transpose(cub(jc), (0, 2, 1))
    want = np.transpose(cub(n), (0, 2, 1))
    out = _interp(yc, n_kc, n_k, B)
    assert np.max(np.abs(out - want)) < 1e-11 * max(1.0, np.max(np.abs(want)))


def test_short_grid_copies():
    B, n_k = 2, 16                                   # n_k - 1 < 3 S: every node
    y = np.random.default_rng(3).standard_normal((B, n_k, 2))
    out = _interp(y, n_k, n_k, B)
    assert np.array_equal(out, y)


def test_solver_psi2_is_interpolated_subgrid():
    from sgl_analytic_amd._lib import PREC_FP32, PSI2_STRIDE as S
    from sgl_analytic_amd.solver import Shape, Solver, cosmo_records
    sh = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256)
    s = Solver(sh, 2, path="binned", precision=PREC_FP32)
    rec = cosmo_records(np.array([[0.674, 0.315, 0.811, 0.0493, 0.965, 3402.0, 0.0, -1.0, 0.0, z]
                                  for z in (0.7, 2.0)]), 8)
    s.run(rec)
    torch.cuda.synchronize()
    assert tuple(s.amp.shape) == (2, 8, s.n_kc, 2) and s.n_kc == 129
    pc = s.psi2c.cpu().numpy()
    pf = s.psi2.cpu().numpy()
    assert np.array_equal(pf[:, ::S, :], pc[:, : len(range(0, 1024, S)), :])
    assert np.all(np.isfinite(pf))
    # small-k limit: Psi_2 -> -k^2 sum_ij C_ij <beta w xi>_i <beta w xi>_j / 2 < 0
    assert np.all(pf[:, 0, 0] < 0)
