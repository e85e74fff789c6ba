"""CPU checks of the C-ABI library: it loads and exports every function
include/ksig_hip.h declares (no compute calls: there is no GPU here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ksig_hip.h")
LIB = os.path.join(ROOT, "ksig_amd", "libksig_hip.so")


def declared_functions():
  text = open(HEADER).read()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  return sorted(set(re.findall(r"\b(ksig_\w+)\s*\(", text)))


def test_header_declares_entry_points():
  names = declared_functions()
  for n in ["ksig_signature_levels", "ksig_signature_levels_gram",
            "ksig_signature_pde", "ksig_signature_pde_gram", "ksig_last_error"]:
    assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_library_exports_every_declared_symbol():
  lib = ctypes.CDLL(LIB)
  for name in declared_functions():
    assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_ctypes_table_matches_header():
  from ksig_amd import _lib
  assert sorted(_lib.SIGNATURES) == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_host_only_entry_points():
  lib = ctypes.CDLL(LIB)
  lib.ksig_version.restype = ctypes.c_char_p
  assert b"gfx950" in lib.ksig_version()
  assert lib.ksig_dp_supported(5, 1) == 1
  assert lib.ksig_dp_supported(8, 2) == 1
  assert lib.ksig_dp_supported(6, 1) == 1
  assert lib.ksig_dp_supported(40, 1) == 0
  # argument validation happens before any device call
  lib.ksig_last_error.restype = ctypes.c_char_p
  rc = lib.ksig_signature_levels_gram(7, None, 0, 1, 1, 5, 1, 1, None, None)
  assert rc == 1001 and b"dtype" in lib.ksig_last_error()
  # any (n_levels, order) runs on caller matrices; a state beyond the LDS of
  # the runtime-levels kernel is refused before launching (dummy pointers
  # are never dereferenced on the host)
  assert lib.ksig_signature_levels_gram(0, None, 0, 4, 4, 40, 1, 1, None, None) == 0
  rc = lib.ksig_signature_levels_gram(0, ctypes.c_void_p(16), 1, 4, 4, 40, 40, 1,
                                      ctypes.c_void_p(16), None)
  assert rc == 1003 and b"LDS" in lib.ksig_last_error()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_workspace_queries_and_adjoint_argument_checks():
  """Host-side sizing of the adjoint workspaces and the validation that runs
  before any device call (dummy non-null pointers are never dereferenced)."""
  from ksig_amd import _lib
  lib = ctypes.CDLL(LIB)
  for name, (res, args) in _lib.SIGNATURES.items():
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = res, args
  # DP adjoint, float32: per pair a [Pc][16] float FIFO, Pc = 128 + 1 -> 144
  assert lib.ksig_signature_levels_bwd_workspace(_lib.F32, 8, 6, 128, 128, 10) == 10 * 144 * 16 * 4
  assert lib.ksig_signature_levels_bwd_workspace(_lib.F32, 8, 6, 20, 30, 1) == 80 * 16 * 4
  assert lib.ksig_signature_levels_bwd_workspace(_lib.F64, 8, 6, 128, 128, 10) == 0
  # PDE adjoint: at dyadic order 0 and 1 per pair one slot per step of the
  # band pass (at least 64 * 2^order + 32) of 4 values and 2 float64;
  # beyond, per pair (refined columns, at least 80) 32-byte (float) slots
  assert lib.ksig_signature_pde_bwd_workspace(_lib.F32, 256, 256, 1, 1, 3) == 3 * 512 * 32
  assert lib.ksig_signature_pde_bwd_workspace(_lib.F64, 40, 10, 0, 0, 2) == 2 * 96 * 48
  assert lib.ksig_signature_pde_bwd_workspace(_lib.F32, 256, 256, 2, 1, 3) == 3 * 1020 * 32
  assert lib.ksig_signature_pde_bwd_workspace(_lib.F64, 40, 10, 3, 0, 2) == 2 * 80 * 64
  p = ctypes.c_void_p(16)  # never dereferenced: validation fails first
  rc = lib.ksig_signature_levels_bwd(
      _lib.F32, p, p, p, 128, p, p, p, 128, 8, p, 4, 0, 1.0, 6, p, p, p, p, p, p, None, 0, None)
  assert rc == 1001 and b"workspace" in lib.ksig_last_error()
  rc = lib.ksig_static_grad_contract(_lib.F32, p, p, p, 4, p, p, p, 4, 0, p, 1, p, p, 1.0, 0.0,
                                     p, p, None)
  assert rc == 1001 and b"d" in lib.ksig_last_error()
  rc = lib.ksig_signature_pde_grad_contract(_lib.F32, p, p, p, 4, p, p, p, 4, 3, p, 1, 7, 1.0,
                                            1, p, p, p, p, None, None)
  assert rc == 1001
  rc = lib.ksig_signature_pde_grad_contract(_lib.F32, p, p, p, 4, p, p, p, 4, 3, p, 1, 0, -1.0,
                                            1, p, p, p, p, None, None)
  assert rc == 1001


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_global_alignment_argument_checks():
  """Validation of the global alignment entry points runs before any device
  call (empty inputs return 0, bad arguments an error code)."""
  lib = ctypes.CDLL(LIB)
  lib.ksig_last_error.restype = ctypes.c_char_p
  f = lib.ksig_global_align_log_gram
  f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                ctypes.c_void_p, ctypes.c_void_p]
  assert f(3, None, 0, 4, 4, None, None) == 1001
  assert f(0, None, -1, 4, 4, None, None) == 1001
  assert f(0, None, 0, 4, 4, None, None) == 0
  assert f(0, None, 2, 4, 4, None, None) == 1001  # null pointers
  g = lib.ksig_global_align_log
  g.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                ctypes.c_void_p, ctypes.c_void_p]
  args = [0, None, None, None, 4, None, None, None, 4, 3, None, 0, 0, 1.0, None, None]
  assert g(*args) == 0
  bad = list(args); bad[12] = 5  # unknown static kind
  assert g(*bad) == 1001
  bad = list(args); bad[13] = -1.0  # non-positive bandwidth
  assert g(*bad) == 1001 and b"positive" in lib.ksig_last_error()
  bad = list(args); bad[9] = 40  # d beyond the built buckets
  assert g(*bad) == 1002


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_random_warping_series_argument_checks():
  """Validation of the random warping series entry points runs before any
  device call."""
  lib = ctypes.CDLL(LIB)
  lib.ksig_last_error.restype = ctypes.c_char_p
  vp, ci, ci64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
  f = lib.ksig_random_warping_series
  f.argtypes = [ci, vp, vp, vp, ci, ci64, ci, vp, vp, ci, ci, vp, ci, vp, vp]
  ok = [0, None, None, None, 10, 0, 3, None, None, 5, 40, None, 1, None, None]
  assert f(*ok) == 0  # no sequences: nothing to do
  bad = list(ok); bad[0] = 4
  assert f(*bad) == 1001 and b"dtype" in lib.ksig_last_error()
  bad = list(ok); bad[6] = 0
  assert f(*bad) == 1001
  bad = list(ok); bad[6] = 33
  assert f(*bad) == 1002
  bad = list(ok); bad[5] = 2  # sequences given, null pointers
  assert f(*bad) == 1001 and b"null" in lib.ksig_last_error()
  g = lib.ksig_random_warping_series_gram
  g.argtypes = [ci, vp, ci64, ci, ci, vp, ci, vp, ci, vp, vp]
  assert g(0, None, 0, 10, 40, None, 5, None, 1, None, None) == 0
  assert g(0, None, 3, -1, 40, None, 5, None, 1, None, None) == 1001
  assert g(0, None, 3, 10, 40, None, 5, None, 1, None, None) == 1001  # null D


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_lowrank_levels_argument_checks():
  lib = ctypes.CDLL(LIB)
  lib.ksig_last_error.restype = ctypes.c_char_p
  f = lib.ksig_lowrank_levels
  f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
  assert f(0, None, 5, 1, 0, 10, 4, None, None) == 0          # no sequences
  assert f(2, None, 5, 1, 3, 10, 4, None, None) == 1001       # dtype
  assert f(0, None, 0, 1, 3, 10, 4, None, None) == 1001       # n_levels
  assert f(0, None, 9, 1, 3, 10, 4, None, None) == 1002       # not built
  assert f(0, None, 5, 3, 3, 10, 4, None, None) == 1001 and b"null" in lib.ksig_last_error()
  assert f(0, None, 5, -1, 3, 10, 4, None, None) == 1001      # order <= 0: all orders


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_lowrank_levels_diag_argument_checks():
  lib = ctypes.CDLL(LIB)
  lib.ksig_last_error.restype = ctypes.c_char_p
  f = lib.ksig_lowrank_levels_diag
  f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
  assert f(0, None, 3, 0, 10, 4, None, None) == 0
  assert f(0, None, 6, 3, 10, 4, None, None) == 1002
  assert f(0, None, 3, 3, 10, 4, None, None) == 1001


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_rff_increments_argument_checks():
  lib = ctypes.CDLL(LIB)
  f = lib.ksig_rff_increments
  f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
  assert f(0, None, 0, 10, 4, 4, 1, None, None) == 0
  assert f(0, None, 3, 1, 4, 4, 1, None, None) == 0     # one point: no increments
  assert f(5, None, 3, 10, 4, 4, 1, None, None) == 1001
  assert f(0, None, 3, 10, 4, 3, 1, None, None) == 1001  # row stride below the row
  assert f(0, None, 3, 10, 4, 8, 1, None, None) == 1001  # null pointers


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_lowrank_levels_diagonal_argument_checks():
  lib = ctypes.CDLL(LIB)
  lib.ksig_last_error.restype = ctypes.c_char_p
  f = lib.ksig_lowrank_levels_diagonal
  f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
  assert f(0, None, 3, 2, 3, 0, 10, 4, None, None) == 0       # no sequences
  assert f(2, None, 3, 1, 2, 3, 10, 4, None, None) == 1001    # dtype
  assert f(0, None, 0, 1, 2, 3, 10, 4, None, None) == 1001    # n_levels
  assert f(0, None, 3, 1, 0, 3, 10, 4, None, None) == 1001    # internal size
  assert f(0, None, 4, 1, 3, 3, 10, 4, None, None) == 1002    # 120 sums: not built
  assert f(0, None, 5, 2, 2, 3, 10, 4, None, None) == 1002    # q=2, 5 levels, order 2
  assert f(0, None, 6, 1, 1, 3, 10, 4, None, None) == 1002    # more than 5 levels
  assert f(0, None, 4, 0, 2, 3, 10, 4, None, None) == 1001 and b"null" in lib.ksig_last_error()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libksig_hip.so not built")
def test_lowrank_levels_sketch_argument_checks():
  lib = ctypes.CDLL(LIB)
  lib.ksig_last_error.restype = ctypes.c_char_p
  f = lib.ksig_lowrank_levels_sketch
  f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                ctypes.c_void_p, ctypes.c_void_p]
  assert f(2, None, 3, 1, 3, 10, 8, None, None, 50, None, None) == 1001   # dtype
  assert f(0, None, 0, 1, 3, 10, 8, None, None, 50, None, None) == 1001   # n_levels
  assert f(0, None, 9, 1, 3, 10, 8, None, None, 50, None, None) == 1002   # > 8 levels
  assert f(0, None, 3, 1, 3, 10, 8, None, None, 0, None, None) == 1001    # n_comp
  assert f(0, None, 3, 1, -1, 10, 8, None, None, 50, None, None) == 1001  # negative shape
  assert f(0, None, 3, 1, 3, 10, 8, None, None, 4096, None, None) == 1002  # > 2048 components
