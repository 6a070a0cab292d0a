"""sgl_lnphi_exact_noalloc (the entry point that allocates nothing) against
sgl_lnphi_exact through the C ABI: two calls in a row, the second on a
smaller input, each bit-identical to the entry point with a caller-owned
work buffer; a request beyond the library array's capacity is rejected."""
import ctypes

import numpy as np
import pytest
import torch

from sgl_analytic_amd import _lib
from sgl_analytic_amd.solver import Shape, cosmo_records

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _tables(L, sh, rec):
    cs = sh.c()
    T = torch.empty(L.sgl_tables_bytes(ctypes.byref(cs), rec.shape[0]) // 8, dtype=torch.float64,
                    device="cuda")
    recd = torch.tensor(rec, dtype=torch.float64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert L.sgl_tables_build(_p(recd), rec.shape[0], ctypes.byref(cs), _p(T), _p(status), None) == 0
    return T, cs


def _both(L, sh, rec):
    B = rec.shape[0]
    T, cs = _tables(L, sh, rec)
    n = B * sh.n_z_max * sh.n_k * 2
    lam_a = torch.full((B * sh.n_k * 2,), np.nan, dtype=torch.float64, device="cuda")
    amp_a = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    work = torch.empty(n, dtype=torch.float64, device="cuda")
    assert L.sgl_lnphi_exact(_p(T), B, ctypes.byref(cs), _p(lam_a), _p(amp_a), _p(work), None) == 0
    lam_b = torch.full_like(lam_a, np.nan)
    amp_b = torch.full_like(amp_a, np.nan)
    assert L.sgl_lnphi_exact_noalloc(_p(T), B, ctypes.byref(cs), _p(lam_b), _p(amp_b), None) == 0
    torch.cuda.synchronize()
    return lam_a.cpu().numpy(), amp_a.cpu().numpy(), lam_b.cpu().numpy(), amp_b.cpu().numpy()


def test_noalloc_matches_the_work_buffer_entry_point():
    L = _lib.load()
    big = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256, filaments=True)
    params = np.array([[0.674, 0.315, 0.811, 0.0493, 0.965, 3402.0, 0.0, -1.0, 0.0, 1.0],
                       [0.7, 0.3, 0.85, 0.045, 0.97, 3402.0, 0.02, -0.9, 0.1, 3.0]])
    rec = cosmo_records(params, np.array([8, 6]))
    print("noalloc call 1: B=2, 8 slices, 1024 k, filaments", flush=True)
    la, aa, lb, ab = _both(L, big, rec)
    assert np.all(np.isfinite(lb)) and np.array_equal(la, lb) and np.array_equal(aa, ab)
    small = Shape(n_z_max=4, n_m=16, n_b=32, n_k=256, n_out=128)
    print("noalloc call 2: B=1, 4 slices, 256 k", flush=True)
    la, aa, lb, ab = _both(L, small, cosmo_records(params[:1], 4))
    assert np.all(np.isfinite(lb)) and np.array_equal(la, lb) and np.array_equal(aa, ab)
    # beyond the array: B * n_z_max * n_k * 2 > 4,194,304 doubles
    huge = Shape(n_z_max=64, n_m=16, n_b=32, n_k=65536, n_out=128)
    T, cs = _tables(L, Shape(n_z_max=4, n_m=16, n_b=32, n_k=256, n_out=128), rec[:1])
    cs_h = huge.c()
    lam = torch.empty(2 * 65536, dtype=torch.float64, device="cuda")
    assert L.sgl_lnphi_exact_noalloc(_p(T), 1, ctypes.byref(cs_h), _p(lam), _p(lam), None) == \
        _lib.SGL_ERR_INVALID
