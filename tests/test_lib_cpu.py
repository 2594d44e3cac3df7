"""The C-ABI library loads and exports every symbol include/dft_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dft_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dft_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("dft_plan_create", "dft_jk", "dft_eri_quartets", "dft_oneelec", "dft_boys", "dft_last_error"):
        assert name in syms


def test_library_exports_all_declared_symbols():
    from dft_amd import _lib
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_error_string():
    from dft_amd import _lib
    lib = _lib.load()
    assert lib.dft_version() >= 100
    # argument validation happens before any HIP call
    rc = lib.dft_plan_create(0, None, None, None, None, None, None, 0, None, None, 0, 0.0, 0.0, 0,
                             ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.DFT_EINVAL
    assert b"nsh" in lib.dft_last_error()


def test_unsupported_angular_momentum_is_rejected():
    import numpy as np
    from dft_amd import _lib
    lib = _lib.load()
    i32 = lambda *v: np.array(v, dtype=np.int32)  # noqa: E731
    f64 = lambda *v: np.array(v, dtype=np.float64)  # noqa: E731
    l, npr, ao = i32(9), i32(1), i32(0)
    xyz, ex, co = f64(0, 0, 0), f64(1.0), f64(1.0)
    rc = lib.dft_plan_create(1, _lib.ptr(l, ctypes.c_int32), _lib.ptr(npr, ctypes.c_int32),
                             _lib.ptr(ao, ctypes.c_int32), _lib.ptr(xyz, ctypes.c_double),
                             _lib.ptr(ex, ctypes.c_double), _lib.ptr(co, ctypes.c_double), 0, None, None, 55,
                             0.0, 0.0, 0, ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.DFT_EUNSUPPORTED
    assert b"angular momentum" in lib.dft_last_error()


def test_eigensolver_rejects_orders_above_its_limit():
    from dft_amd import _lib
    lib = _lib.load()
    rc = lib.dft_syevj(129, 1, None, None, None, 1, None, None)
    assert rc == _lib.DFT_EUNSUPPORTED
    assert b"above 128" in lib.dft_last_error()
    # a negative or NaN off-diagonal tolerance is an argument error
    for tol in (-1e-9, float("nan")):
        rc = lib.dft_syevj_tol(16, 1, None, None, None, 1, None, tol, None)
        assert rc == _lib.DFT_EINVAL
        assert b"tol" in lib.dft_last_error()


def test_scf_dense_products_reject_bad_shapes():
    """dft_scf_gemm / _fock_residual / _orbital_matrix check orders (1..128)
    and leading dimensions before touching the device (placeholder
    pointers: nothing is read)."""
    from dft_amd import _lib
    lib = _lib.load()
    f = ctypes.c_void_p(0x1000)
    assert lib.dft_scf_gemm(129, 8, 8, f, 8, 0, f, 8, 0, f, 8, None, None, None) == _lib.DFT_EUNSUPPORTED
    assert lib.dft_scf_gemm(8, 8, 0, f, 8, 0, f, 8, 0, f, 8, None, None, None) == _lib.DFT_EUNSUPPORTED
    assert lib.dft_scf_gemm(8, 8, 8, f, 7, 0, f, 8, 0, f, 8, None, None, None) == _lib.DFT_EINVAL
    assert b"leading dimension" in lib.dft_last_error()
    assert lib.dft_scf_gemm(8, 8, 8, f, 8, 0, f, 8, 0, f, 8, f, None, None) == _lib.DFT_EINVAL
    assert lib.dft_scf_fock_residual(0, f, f, f, f, f, f, None) == _lib.DFT_EUNSUPPORTED
    assert lib.dft_scf_orbital_matrix(129, f, f, f, f, f, None) == _lib.DFT_EUNSUPPORTED
    assert lib.dft_scf_orbital_matrix(8, f, None, f, f, f, None) == _lib.DFT_EINVAL


def test_eigensolver_workspace_size_and_check():
    """dft_syevj_ws_bytes: a 1 KiB header and 40 sweeps of (m - 1) rounds of
    m / 2 rotations of 16 bytes (m = n rounded up to even); dft_syevj_ws
    rejects a workspace below batch times that before it touches the device
    (placeholder pointers: nothing is read)."""
    from dft_amd import _lib
    lib = _lib.load()
    for n in (1, 2, 7, 112, 128):
        m = n + (n & 1)
        assert lib.dft_syevj_ws_bytes(n) == 1024 + 40 * (m - 1) * (m // 2) * 16
    assert lib.dft_syevj_ws_bytes(112) == 3_979_264
    for n in (0, 129, -3):
        assert lib.dft_syevj_ws_bytes(n) == 0
    fake = ctypes.c_void_p(0x1000)
    need = 3 * lib.dft_syevj_ws_bytes(40)
    rc = lib.dft_syevj_ws(40, 3, fake, fake, fake, 1, None, 0.0, fake, need - 1, None)
    assert rc == _lib.DFT_EINVAL
    assert b"work_bytes" in lib.dft_last_error()
    rc = lib.dft_syevj_ws(129, 1, fake, fake, fake, 1, None, 0.0, fake, need, None)
    assert rc == _lib.DFT_EUNSUPPORTED


def test_python_constants_match_the_header():
    from dft_amd import _lib
    text = open(os.path.join(ROOT, "include", "dft_hip.h")).read()
    for name in ("DFT_PLAN_UHF", "DFT_PLAN_NOSPLIT", "DFT_PLAN_SERIAL", "DFT_PLAN_STATIC_WORKSPACE", "DFT_PLAN_MFMA"):
        m = re.search(r"#define\s+%s\s+(\d+)" % name, text)
        assert m and int(m.group(1)) == getattr(_lib, name), name
    m = re.search(r"#define\s+DFT_STATIC_SCRATCH_BYTES\s+\((\d+)LL << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.DFT_STATIC_SCRATCH_BYTES
