"""ctypes binding of libksig_hip.so (the C ABI declared in include/ksig_hip.h).

The HIP library is the only compute path of this package: if it is missing,
or no GPU is visible, every compute entry point raises instead of falling back.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

_LIB_PATH = Path(__file__).resolve().parent / "libksig_hip.so"
_lib = None

F32, F64 = 0, 1
STATIC_RBF, STATIC_LINEAR = 0, 1
GRAM_AUTO, GRAM_VALU, GRAM_MFMA_F32, GRAM_MFMA_BF16X3 = -1, 0, 1, 2
GRAM_MODES = {'auto': GRAM_AUTO, 'valu': GRAM_VALU, 'fp32': GRAM_MFMA_F32,
              'bf16x3': GRAM_MFMA_BF16X3}

_c_int, _c_i64, _c_dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
_c_size = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/ksig_hip.h
SIGNATURES = {
    "ksig_version": (ctypes.c_char_p, []),
    "ksig_last_error": (ctypes.c_char_p, []),
    "ksig_dp_supported": (_c_int, [_c_int, _c_int]),
    "ksig_signature_levels": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _c_int, _c_int, _c_int, _vp, _vp]),
    "ksig_stream_slots": (_c_int, [
        _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "ksig_signature_levels_batched": (_c_int, [
        _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_i64, _c_int,
        _c_dbl, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp, _c_size, _vp, _vp]),
    "ksig_signature_levels_batched_noalloc": (_c_int, [
        _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_i64, _c_int,
        _c_dbl, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp, _vp]),
    "ksig_noalloc_scratch_bytes": (_c_size, []),
    "ksig_signature_levels_fwd_save": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ksig_signature_levels_fwd_save_workspace": (_c_size, [
        _c_int, _c_int, _c_int, _c_int, _c_int, _c_i64]),
    "ksig_signature_levels_bwd": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ksig_signature_levels_bwd_workspace": (_c_size, [
        _c_int, _c_int, _c_int, _c_int, _c_int, _c_i64]),
    "ksig_signature_pde_grad_contract": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_i64,
        _c_int, _c_dbl, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ksig_static_grad_contract": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_i64,
        _vp, _vp, _c_dbl, _c_dbl, _vp, _vp, _vp]),
    "ksig_signature_levels_gram": (_c_int, [
        _c_int, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp]),
    "ksig_signature_levels_dots": (_c_int, [
        _c_int, _vp, _vp, _c_i64, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
        _c_i64, _c_int, _c_int, _c_int, _vp, _vp]),
    "ksig_signature_pde": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp]),
    "ksig_signature_pde_fwd_edges": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "ksig_signature_pde_bwd_edges": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "ksig_signature_pde_bwd_workspace": (_c_size, [
        _c_int, _c_int, _c_int, _c_int, _c_int, _c_i64]),
    "ksig_signature_pde_bwd_fused": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size,
        _vp]),
    "ksig_signature_pde_bwd_fused_workspace": (_c_size, [
        _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_i64]),
    "ksig_signature_pde_gram": (_c_int, [
        _c_int, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _vp, _vp]),
    "ksig_global_align_log": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _c_int, _vp,
        _c_i64, _c_int, _c_dbl, _vp, _vp]),
    "ksig_global_align_log_gram": (_c_int, [
        _c_int, _vp, _c_i64, _c_int, _c_int, _vp, _vp]),
    "ksig_random_warping_series": (_c_int, [
        _c_int, _vp, _vp, _vp, _c_int, _c_i64, _c_int, _vp, _vp, _c_int, _c_int, _vp,
        _c_int, _vp, _vp]),
    "ksig_random_warping_series_gram": (_c_int, [
        _c_int, _vp, _c_i64, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp]),
    "ksig_lowrank_levels": (_c_int, [_c_int, _vp, _c_int, _c_int, _c_i64, _c_int, _c_int, _vp,
                                     _vp]),
    "ksig_lowrank_levels_diag": (_c_int, [_c_int, _vp, _c_int, _c_i64, _c_int, _c_int, _vp,
                                          _vp]),
    "ksig_rff_increments": (_c_int, [_c_int, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _vp,
                                     _vp]),
    "ksig_lowrank_levels_diagonal": (_c_int, [_c_int, _vp, _c_int, _c_int, _c_int, _c_i64,
                                              _c_int, _c_int, _vp, _vp]),
    "ksig_lowrank_levels_sketch": (_c_int, [_c_int, _vp, _c_int, _c_int, _c_i64, _c_int, _c_int,
                                            _vp, _vp, _c_int, _vp, _vp]),
}


class KsigError(RuntimeError):
  """A C-ABI call returned a non-zero status."""


def library_path() -> Path:
  return Path(os.environ.get("KSIG_HIP_LIB", _LIB_PATH))


def load(require_gpu: bool = True):
  """Load libksig_hip.so once and bind the C-ABI symbols."""
  global _lib
  if require_gpu and not torch.cuda.is_available():
    raise RuntimeError(
        "ksig_amd computes on an AMD GPU through libksig_hip.so; no GPU is "
        "visible to torch (torch.cuda.is_available() is False)")
  if _lib is None:
    path = library_path()
    if not path.exists():
      raise RuntimeError(
          f"{path} is missing: build it with `python -m ksig_amd.build` "
          "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
      fn = getattr(lib, name)
      fn.restype = res
      fn.argtypes = args
    _lib = lib
  return _lib


def check(status: int) -> None:
  if status != 0:
    msg = _lib.ksig_last_error().decode(errors="replace") if _lib else ""
    if status in (1001,):
      raise ValueError(msg or f"ksig_hip error {status}")
    if status in (1002,):
      raise NotImplementedError(msg or f"ksig_hip error {status}")
    raise KsigError(f"ksig_hip error {status}: {msg}")


def stream_handle(device: torch.device) -> int:
  return torch.cuda.current_stream(device).cuda_stream


def dtype_code(dtype: torch.dtype) -> int:
  if dtype == torch.float32:
    return F32
  if dtype == torch.float64:
    return F64
  raise TypeError(f"unsupported dtype {dtype} (float32 or float64)")
