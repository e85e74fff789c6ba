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
  rc = lib.ksig_signature_levels_gram(0, None, 0, 4, 4, 40, 1, 1, None, None)
  assert rc == 1002
