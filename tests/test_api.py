"""CPU checks of the host-side API mirror (parameters, validation, errors)."""

import numpy as np
import pytest
import torch

import ksig_amd as ksig


def test_constructor_params_and_order_normalisation():
  k = ksig.kernels.SignatureKernel(n_levels=4, order=0)
  assert k.order == 4
  k = ksig.kernels.SignatureKernel(n_levels=4, order=7)
  assert k.order == 4
  k = ksig.kernels.SignatureKernel(n_levels=4, order=2)
  assert k.order == 2
  p = k.get_params()
  assert p["n_levels"] == 4 and p["static_kernel__bandwidth"] == 1.0
  pde = ksig.kernels.SignaturePDEKernel(dyadic_order=2)
  assert pde.get_params()["dyadic_order"] == 2


def test_positive_parameter_checks():
  with pytest.raises(ValueError):
    ksig.kernels.SignatureKernel(n_levels=0)
  with pytest.raises(ValueError):
    ksig.kernels.SignatureKernel(n_features=-1)
  with pytest.raises(ValueError):
    ksig.static.kernels.RBFKernel(bandwidth=0.)
  with pytest.raises(ValueError):
    ksig.static.kernels.LinearKernel(scale=-1.)
  with pytest.raises(ValueError):
    ksig.kernels.SignaturePDEKernel(dyadic_order=-1)


def test_validate_data_shapes():
  k = ksig.kernels.SignatureKernel(n_features=3)
  X = np.zeros((4, 12))
  assert k._validate_data(X).shape == (4, 4, 3)
  with pytest.raises(ValueError):
    k._validate_data(np.zeros((4, 5, 2)))
  with pytest.raises(ValueError):
    ksig.kernels.SignatureKernel()._validate_data(np.zeros((2, 3, 4, 5)))
  with pytest.warns(UserWarning):
    assert ksig.kernels.SignatureKernel()._validate_data(np.zeros((4, 6))).shape == (4, 6, 1)


def test_pde_rejects_bad_ndim_before_compute():
  # ndim check of ksig/algorithms.py:417-418 comes before any device work
  if torch.cuda.is_available():
    pytest.skip("GPU present; covered by GPU tests")
  with pytest.raises(RuntimeError):
    ksig.algorithms.signature_kern_pde(np.zeros((2, 3)))


def test_compute_without_gpu_fails_loudly():
  if torch.cuda.is_available():
    pytest.skip("GPU present")
  with pytest.raises(RuntimeError, match="GPU"):
    ksig.kernels.SignatureKernel()(np.zeros((2, 5, 3)))
