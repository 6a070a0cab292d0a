"""The C-ABI library loads on a host without a GPU and exports every entry
point include/sgl_amd.h declares; host-only size queries work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sgl_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(sgl_\w+)\(", text, flags=re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "sgl_tables_build" in names and "sgl_invert_gemm" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from sgl_analytic_amd import _lib
    L = _lib.load()
    for name in declared():
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) <= set(declared())


def test_size_queries_and_rejections():
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import Shape
    L = _lib.load()
    sh = Shape().c()
    n = L.sgl_tables_bytes(ctypes.byref(sh), 4096)
    assert n > 4096 * 64 * 256 * 8 * 8
    assert L.sgl_hist_bytes(ctypes.byref(sh), 2) > 0
    assert L.sgl_filon_bytes(ctypes.byref(sh), _lib.PREC_FP32) == 2048 * 16448 * 4
    assert L.sgl_filon_bytes(ctypes.byref(sh), _lib.PREC_BF16X3) == 2 * 2048 * 16448 * 2
    assert L.sgl_phi_bytes(ctypes.byref(sh), 3, _lib.PREC_FP64) == 3 * 8193 * 16
    bad = Shape(n_z_max=200).c()                 # above the 96-slice limit
    assert L.sgl_tables_bytes(ctypes.byref(bad), 1) == 0
    assert L.sgl_tables_bytes(ctypes.byref(sh), 0) == 0
    # invalid arguments are rejected on the host before any launch
    rc = L.sgl_tables_build(None, 1, ctypes.byref(sh), None, None, None)
    assert rc == _lib.SGL_ERR_INVALID
    rc = L.sgl_invert_gemm(None, None, 1, ctypes.byref(sh), 7, None, None)
    assert rc == _lib.SGL_ERR_INVALID
    assert L.sgl_version().startswith(b"sgl_amd")


def test_psi2_subgrid_node_count():
    """Two-point sub-grid: node jc is k node jc * 8; the last sub-node covers
    k node n_k - 1 (it may lie past k_hi); short grids keep every node."""
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import Shape
    L = _lib.load()
    S = _lib.PSI2_STRIDE
    for n_k in (8192, 8193, 1024, 512, 25, 24, 2):
        sh = Shape(n_k=n_k).c()
        n_kc = L.sgl_psi2_nodes(ctypes.byref(sh))
        if n_k - 1 >= 3 * S:
            assert (n_kc - 1) * S >= n_k - 1 > (n_kc - 2) * S, (n_k, n_kc)
        else:
            assert n_kc == n_k
    assert L.sgl_psi2_nodes(ctypes.byref(Shape().c())) == 1025
    assert L.sgl_psi2_nodes(None) == 0


def test_vp_workspace_size():
    """sgl_vp_bytes: eta and target (2B), wk and the rate table per group,
    one 48 x 48 transition per slice, ytot and the 48 rate-weighted sums per
    (slice, k node), each sub-array padded to 32 doubles; groups are the mass
    rows, twice that with filaments; invalid shapes give 0."""
    import ctypes
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import Shape
    L = _lib.load()

    def expect(sh, B):
        pad = lambda n: (n + 31) // 32 * 32
        nz, G = B * sh.n_z_max, sh.n_m * (2 if sh.filaments else 1)
        return 8 * (pad(2 * B) + pad(nz * G) + pad(nz * 48 * 48) + pad(nz * G * 3)
                    + pad(nz * sh.n_k * 2) + pad(nz * 48 * sh.n_k * 2))

    for sh, B in ((Shape(), 1), (Shape(filaments=True), 3), (Shape(n_z_max=7, n_m=33, n_k=999), 5)):
        assert L.sgl_vp_bytes(ctypes.byref(sh.c()), B) == expect(sh, B)
    assert L.sgl_vp_bytes(ctypes.byref(Shape(n_z_max=200).c()), 1) == 0
    assert L.sgl_vp_bytes(ctypes.byref(Shape().c()), 0) == 0
