"""Kernel-level GPU checks: the MFMA complex GEMM and the blocked LU of
librefeff_fms.so against numpy / the oracle's LAPACK-style LU, in FP32 and FP64."""
import ctypes as C

import numpy as np
import pytest

import refeff_amd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _lib():
    from refeff_amd import _lib as rl
    L = rl.test_lib()
    L.refeff_fms_debug_lu.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]
    L.refeff_fms_debug_gemm.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("use_double", [0, 1])
@pytest.mark.parametrize("m,n,k", [(32, 32, 32), (64, 96, 32), (160, 96, 64), (224, 288, 32)])
def test_gemm_sub(use_double, m, n, k):
    _gemm_case(use_double, m, n, k)


@pytest.mark.parametrize("kind", [3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("m,n,k", [(256, 256, 128), (320, 448, 64), (130, 260, 32), (96, 16, 32), (2880, 16, 32)])
def test_gemm_v2(kind, m, n, k):
    """64x64-tile double-buffered FP32 kernel of the LU updates: BK 32 / 16, with and
    without the XCD-aware tile order (kinds 3, 4, 5, 6), and the 3M form (kind 7; kind 8 with
    the C tile prefetched)."""
    _gemm_case(kind, m, n, k)


@pytest.mark.parametrize("kind,tol", [(12, 6e-5), (13, 2e-5), (14, 2e-5), (15, 2e-5)])
@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (256, 256, 128), (320, 448, 64), (1024, 512, 256), (192, 640, 96)])
def test_gemm_split_bf16(kind, tol, m, n, k):
    """Trailing update on the bf16 MFMA from split operands: two parts (kind 12,
    relative error of order 2^-17 per product) and three parts (kind 13, FP32-like);
    on the f16 MFMA from two parts with power-of-two row / column scales (kind 14,
    3M form, and kind 15, four-product form: FP32-like, 128 x 128 tiles, so the
    64-multiple shapes leave partial tiles)."""
    _gemm_case(kind, m, n, k, tol)


@pytest.mark.parametrize("m,n,k", [(3072, 3072, 256), (3008, 3136, 128), (2944, 3072, 192)])
def test_gemm_split_fp16_large(m, n, k):
    """The four-product split-fp16 update over more tiles than the GPU has CUs
    (576 and more), with partial tiles at the right and bottom edges and the
    half of each wave's C block that is loaded during the last two K steps."""
    _gemm_case(15, m, n, k, 2e-5)


@pytest.mark.parametrize("kind", [14, 15])
def test_gemm_split_fp16_keeps_scaled_rows_and_columns(kind):
    """Rows of A and columns of B scaled over six decades: with a power-of-two
    scale per row and per column every element of C - A B keeps an FP32-like
    error relative to its own row and column scale, not just to the largest."""
    m, n, k = 256, 320, 256
    print(f"[gpu-kernels] gemm kind={kind} scaled rows / columns {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(7)
    r = 10.0 ** rng.uniform(-3, 3, m)
    cs = 10.0 ** rng.uniform(-3, 3, n)
    a, b = crand(rng, m, k) * r[:, None], crand(rng, k, n) * cs[None, :]
    c = crand(rng, m, n) * r[:, None] * cs[None, :]
    ref = c - a @ b
    out = np.ascontiguousarray(c.copy())
    rc = _lib().refeff_fms_debug_gemm(kind, m, n, k, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), _p(out))
    assert rc == 0
    err = np.abs(out - ref) / (r[:, None] * cs[None, :])
    scale = np.abs(ref / (r[:, None] * cs[None, :])).max()
    assert err.max() <= 2e-5 * scale, err.max() / scale


def _gemm_case(use_double, m, n, k, tol=None):
    print(f"[gpu-kernels] gemm double={use_double} {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m + n + k)
    a, b, c = crand(rng, m, k), crand(rng, k, n), crand(rng, m, n)
    ref = c - a @ b
    out = np.ascontiguousarray(c.copy())
    rc = _lib().refeff_fms_debug_gemm(use_double, m, n, k, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)),
                                      _p(out))
    assert rc == 0
    if tol is None:
        tol = 1e-12 if use_double == 1 else 2e-5
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("n", [64, 192])
def test_lu_double_matches_lapack_style_lu(n):
    print(f"[gpu-kernels] lu double n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = crand(rng, n, n)
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, 1, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info)) == 0
    assert info.value == 0
    ref_lu, ref_piv = orc.zgetrf(a)
    assert np.array_equal(piv, ref_piv)
    assert np.abs(work - ref_lu).max() <= 1e-11 * np.abs(ref_lu).max()


@pytest.mark.parametrize("n", [2112])
def test_lu_double_large_matches_lapack(n):
    """Panels of more than 2048 rows (several register rows per thread)."""
    import scipy.linalg.lapack as lapack
    print(f"[gpu-kernels] lu double n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = crand(rng, n, n)
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, 1, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info)) == 0
    assert info.value == 0
    ref_lu, ref_piv, ref_info = lapack.zgetrf(a)
    assert ref_info == 0
    assert np.array_equal(piv, ref_piv)
    assert np.abs(work - ref_lu).max() <= 1e-10 * np.abs(ref_lu).max()


@pytest.mark.parametrize("panel", ["calu", "gepp"])
@pytest.mark.parametrize("n", [64, 256, 1152, 4160])
def test_lu_float_reconstructs_matrix(n, panel, monkeypatch):
    """FP32 LU with tournament-pivoting panels (default; 1152 rows take one
    selection level of 256-row workgroups, 4160 rows a second of 512-row ones)
    and with the row-by-row partial-pivoting panels."""
    monkeypatch.setenv("REFEFF_FMS_PANEL", panel)
    print(f"[gpu-kernels] lu float n={n} panel={panel}", flush=True)
    rng = np.random.default_rng(100 + n)
    a = crand(rng, n, n)
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, 0, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info)) == 0
    assert info.value == 0
    lo = np.tril(work, -1) + np.eye(n)
    up = np.triu(work)
    pa = a.copy()
    for i, p in enumerate(piv):
        pa[[i, p]] = pa[[p, i]]
    assert np.abs(lo @ up - pa).max() <= 1e-4 * np.sqrt(n / 64) * np.abs(a).max()


@pytest.mark.parametrize("use_double,block,panel", [(1, 160, "gepp"), (1, 512, "gepp"), (0, 160, "calu"),
                                                    (0, 512, "calu"), (0, 224, "gepp")])
def test_lu_recursive_halving(use_double, block, panel, monkeypatch):
    """Outer blocks factored by recursive halving (160 = 64 + 96, then 32 + 64;
    512 = 256 + 256 down to 32; 224 = 96 + 128), with each panel's interchanges
    applied to the whole block at once and the solves and updates deferred: in
    FP64 the same pivots as LAPACK zgetrf and the same factors to rounding, in
    FP32 factors that reconstruct P A.  1088 rows leave a partial last block."""
    import scipy.linalg.lapack as lapack
    monkeypatch.setenv("REFEFF_FMS_LU_SPLIT", "1")
    monkeypatch.setenv("REFEFF_FMS_LU_BLOCK", str(block))
    monkeypatch.setenv("REFEFF_FMS_PANEL", panel)
    n = 1088
    print(f"[gpu-kernels] lu recursive halving n={n} block={block} double={use_double} panel={panel}", flush=True)
    rng = np.random.default_rng(block + use_double)
    a = crand(rng, n, n)
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, use_double, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)),
                                      C.byref(info)) == 0
    assert info.value == 0
    if use_double:
        ref_lu, ref_piv, ref_info = lapack.zgetrf(a)
        assert ref_info == 0
        assert np.array_equal(piv, ref_piv)
        assert np.abs(work - ref_lu).max() <= 1e-10 * np.abs(ref_lu).max()
    else:
        lo = np.tril(work, -1) + np.eye(n)
        up = np.triu(work)
        pa = a.copy()
        for i, p in enumerate(piv):
            pa[[i, p]] = pa[[p, i]]
        assert np.abs(lo @ up - pa).max() <= 1e-4 * np.sqrt(n / 64) * np.abs(a).max()


@pytest.mark.parametrize("n,use_double,panel", [(9792, 0, "calu"), (9792, 0, "gepp"), (8256, 1, "gepp")])
def test_lu_tall_panels_solve_with_small_backward_error(n, use_double, panel, monkeypatch):
    """Panels of 8192-10240 rows (FP32: two tournament selection levels, or register
    tiles of 2 columns; FP64: register tiles of 1 column): the factors solve
    A x = b with a backward error at the working precision."""
    import scipy.linalg as sla
    monkeypatch.setenv("REFEFF_FMS_PANEL", panel)
    print(f"[gpu-kernels] lu tall panels n={n} double={use_double} panel={panel}", flush=True)
    rng = np.random.default_rng(n)
    a = crand(rng, n, n) / np.sqrt(n) + 2.0 * np.eye(n)
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, use_double, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)),
                                      C.byref(info)) == 0
    assert info.value == 0
    assert np.all((piv >= np.arange(n)) & (piv < n))
    b = crand(rng, n, 2)
    x = sla.lu_solve((work, piv), b, check_finite=False)
    berr = np.linalg.norm(a @ x - b) / (np.linalg.norm(a, 2 if n < 0 else "fro") * np.linalg.norm(x) + np.linalg.norm(b))
    assert berr <= (1e-13 if use_double else 5e-6), berr


def test_lu_reports_exact_singularity():
    print("[gpu-kernels] lu singular", flush=True)
    n = 64
    rng = np.random.default_rng(5)
    a = crand(rng, n, n)
    a[:, 7] = 0.0
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, 1, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info)) == 0
    assert info.value == 8  # first zero pivot, one-based, as LAPACK info


def test_lu_float_tournament_reports_singularity_and_top_row_pivots(monkeypatch):
    """Tournament-pivoting panels: exactly zero columns give LAPACK's info (first
    zero pivot, one-based) and pivots drawn from the panel's own top rows and from
    far below both land as valid interchanges."""
    monkeypatch.setenv("REFEFF_FMS_PANEL", "calu")
    print("[gpu-kernels] lu float tournament singular / top-row pivots", flush=True)
    n = 1088
    rng = np.random.default_rng(9)
    a = crand(rng, n, n)
    a[:16, :32] *= 50.0       # half the first panel's pivots come from its own top rows
    a[900:916, :32] *= 1e-3    # (never pivots of the first panel, so they stay in place)
    a[900:916, 32:64] *= 50.0  # the second panel's from a later selection chunk
    work = np.ascontiguousarray(a.copy())
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    assert _lib().refeff_fms_debug_lu(n, 0, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info)) == 0
    assert info.value == 0
    assert np.all((piv >= np.arange(n)) & (piv < n))
    assert set(range(900, 916)) <= set(piv[32:64].tolist())
    lo = np.tril(work, -1) + np.eye(n)
    up = np.triu(work)
    pa = a.copy()
    for i, p in enumerate(piv):
        pa[[i, p]] = pa[[p, i]]
    assert np.abs(lo @ up - pa).max() <= 1e-4 * np.sqrt(n / 64) * np.abs(a).max()
    s = a.copy()
    s[:, 7] = 0.0
    s[:, 600] = 0.0
    work = np.ascontiguousarray(s)
    assert _lib().refeff_fms_debug_lu(n, 0, _p(work), piv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info)) == 0
    assert info.value == 8
