"""CPU checks of the reference-name package ``ksig`` and of the host-side
counterparts it exposes: the import lines of the reference's experiment
harness (experiments/experiment.py:7-16), the static kernels beyond RBF and
linear against closed forms, the preprocessing transformers against NumPy
restatements of ksig/preprocessing.py, and PrecomputedFeatureLinSVC on a
host feature map.  The torch code paths run on CPU tensors here (they are
device-agnostic); the signature kernels themselves need the GPU."""

import numpy as np
import pytest
import torch
from sklearn.base import clone


def test_reference_import_lines():
  import ksig
  from ksig.static.kernels import Kernel
  from ksig.models import PrecomputedFeatureLinSVC, PrecomputedKernelSVC
  from ksig.utils import _EPS, ArrayOnCPUOrGPU, ArrayOnGPU
  import ksig_amd
  assert _EPS == 1e-12
  assert ksig.kernels is ksig_amd.kernels
  assert ksig.static.kernels is ksig_amd.static.kernels
  assert Kernel is ksig_amd.static.kernels.Kernel
  for name in ("SignatureKernel", "SignaturePDEKernel", "SignatureFeatures",
               "GlobalAlignmentKernel", "RandomWarpingSeries"):
    assert hasattr(ksig.kernels, name)
  for name in ("PolynomialKernel", "Matern12Kernel", "Matern32Kernel", "Matern52Kernel",
               "RationalQuadraticKernel", "RBFKernel", "LinearKernel"):
    assert hasattr(ksig.static.kernels, name)
  for name in ("RandomFourierFeatures", "RandomFourierFeatures1D", "NystroemFeatures"):
    assert hasattr(ksig.static.features, name)
  for name in ("signature_kern", "signature_kern_pde"):
    assert hasattr(ksig.algorithms, name)
  assert hasattr(ksig.preprocessing, "SequenceAugmentor")
  assert isinstance(np.zeros(1), ArrayOnCPUOrGPU.__args__)
  assert ArrayOnGPU is torch.Tensor
  assert issubclass(PrecomputedFeatureLinSVC, PrecomputedKernelSVC)


def _d2(A, B):
  return ((A[:, None, :] - B[None, :, :]) ** 2).sum(-1)


@pytest.mark.parametrize("name", ["poly", "m12", "m32", "m52", "rq"])
def test_static_kernels_closed_forms(name):
  from ksig.static import kernels as sk
  rng = np.random.default_rng(0)
  A = rng.standard_normal((7, 4))
  B = rng.standard_normal((5, 4))
  s = 1.7
  D = np.sqrt(np.maximum(_d2(A, B), 1e-12))
  if name == "poly":
    k, ref = sk.PolynomialKernel(degree=3., gamma=0.5, scale=2.), 2. * (A @ B.T + 0.5) ** 3
  elif name == "m12":
    k, ref = sk.Matern12Kernel(s), np.exp(-D / s)
  elif name == "m32":
    r = np.sqrt(3.) * D / s
    k, ref = sk.Matern32Kernel(s), (1 + r) * np.exp(-r)
  elif name == "m52":
    r2 = 5 * _d2(A, B) / s ** 2
    r = np.sqrt(np.maximum(r2, 1e-12))
    k, ref = sk.Matern52Kernel(s), (1 + r + r2 / 3) * np.exp(-r)
  else:
    k, ref = sk.RationalQuadraticKernel(s, alpha=0.7), (1 + _d2(A, B) / (2 * 0.7 * s * s)) ** -0.7
  got = k._K(torch.as_tensor(A), torch.as_tensor(B)).numpy()
  np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
  # symmetric call (the reference's Matern12/32 fail there, see the class docstring)
  sym = k._K(torch.as_tensor(A)).numpy()
  np.testing.assert_allclose(sym, sym.T, rtol=0, atol=1e-12)
  assert clone(k).get_params() == k.get_params()
  if name != "poly":
    np.testing.assert_array_equal(k._Kdiag(torch.as_tensor(A)).numpy(), 1)


def test_static_kernel_parameter_checks():
  from ksig.static import kernels as sk
  with pytest.raises(ValueError):
    sk.PolynomialKernel(degree=0)
  with pytest.raises(ValueError):
    sk.RationalQuadraticKernel(alpha=-1.)
  with pytest.raises(ValueError):
    sk.Matern52Kernel(bandwidth=0.)


def _aug_numpy(X, add_time, lead_lag, basepoint, normalize, max_time, max_len):
  """ksig/preprocessing.py:105-138 restated with NumPy (on a copy)."""
  X = X.copy()
  if normalize:
    X = X / X.max()
  if lead_lag:
    X = np.repeat(X, 2, axis=1)
    X = np.concatenate((X[:, 1:], X[:, :-1]), axis=-1)
  if add_time and max_time > 1e-6:
    t = np.linspace(0., max_time, X.shape[1])
    X = np.concatenate((np.tile(t[None, :, None], [X.shape[0], 1, 1]), X), axis=-1)
  if basepoint:
    X = np.concatenate((np.zeros_like(X[:, :1]), X), axis=1)
  if max_len is not None and X.shape[1] > max_len:
    cur, tgt = np.linspace(0, 1, X.shape[1]), np.linspace(0, 1, max_len)
    X = np.apply_along_axis(lambda v: np.interp(tgt, cur, v), 1, X)
  return X


@pytest.mark.parametrize("flags", [(True, True, True, True, 1., None),
                                   (False, True, False, False, 1., None),
                                   (True, False, True, True, 10., 9),
                                   (True, True, False, True, 1e-9, 12)])
def test_sequence_augmentor(flags):
  from ksig.preprocessing import SequenceAugmentor
  rng = np.random.default_rng(1)
  X = rng.random((3, 8, 2)) + 0.1
  ref = _aug_numpy(X, *flags)
  kw = dict(zip(("add_time", "lead_lag", "basepoint", "normalize", "max_time", "max_len"),
                flags))
  got = SequenceAugmentor(**kw).fit(X.copy()).transform(X.copy())
  np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14)
  Xt = torch.as_tensor(X.copy())
  gott = SequenceAugmentor(**kw).fit(Xt).transform(Xt.clone())
  np.testing.assert_allclose(gott.numpy(), ref, rtol=0, atol=1e-12)
  # lead-lag by hand on one channel: (x0,x0,x1,x1,...) leads, lags one behind
  if flags[1] and not flags[0] and not flags[2]:
    x = X[0, :, 0] / (X.max() if flags[3] else 1.0)
    np.testing.assert_allclose(got[0, :, 0], np.repeat(x, 2)[1:])
    np.testing.assert_allclose(got[0, :, 2], np.repeat(x, 2)[:-1])


def test_sequence_tabulator():
  from ksig.preprocessing import SequenceTabulator
  rng = np.random.default_rng(2)
  seqs = [rng.standard_normal((L, 3)) for L in (5, 9, 7)]
  seqs[1][3, 1] = np.nan
  tab = SequenceTabulator().fit(seqs)
  assert tab.max_len_ == 9
  out = tab.transform(seqs)
  assert out.shape == (3, 9, 3)
  # row with a NaN dropped, then each channel interpolated to 9 points
  s1 = seqs[1][~np.isnan(seqs[1]).any(-1)]
  for c in range(3):
    np.testing.assert_allclose(out[1, :, c], np.interp(np.linspace(0, 1, 9),
                                                       np.linspace(0, 1, 8), s1[:, c]))
  np.testing.assert_allclose(out[0, 0], seqs[0][0])
  np.testing.assert_allclose(out[0, -1], seqs[0][-1])
  outt = SequenceTabulator().fit(seqs).transform([torch.as_tensor(s) for s in seqs])
  np.testing.assert_allclose(outt.numpy(), out, rtol=0, atol=1e-12)
  # equal lengths without NaNs pass through unchanged
  same = [rng.standard_normal((6, 2)) for _ in range(4)]
  np.testing.assert_array_equal(SequenceTabulator().fit(same).transform(same), np.stack(same))
  assert SequenceTabulator(max_len=4).fit(same).transform(same).shape == (4, 4, 2)


class HostFeatures:
  """A feature map that is not one of this library's (host NumPy)."""

  def fit(self, X):
    return self

  def transform(self, X, return_on_gpu=False):
    return np.concatenate([X.mean(1), X.std(1)], -1)


def test_feature_linsvc_fit_predict_and_params():
  from ksig.models import PrecomputedFeatureLinSVC
  rng = np.random.default_rng(3)
  n = 40
  y = np.arange(n) % 2
  X = rng.standard_normal((n, 12, 3)) + np.where(y == 0, -1.0, 1.0)[:, None, None]
  m = PrecomputedFeatureLinSVC(HostFeatures(), svc_hparams={}, batch_size=7)
  assert m.svc_hparams["dual"] is False and m.svc_hparams["fit_intercept"] is False
  m.fit(X, y)
  assert (m.predict(X) == y).mean() > 0.95
  F = m._precompute_model_inputs()
  np.testing.assert_allclose(F, HostFeatures().transform(X))
  g = PrecomputedFeatureLinSVC(HostFeatures(), svc_hparams={}, svc_grid={"C": [0.1, 1.0]},
                               cv=3, n_jobs=1).fit(X, y)
  assert (g.predict(X) == y).mean() > 0.95
  assert "on_gpu" in clone(m).get_params()
  assert PrecomputedFeatureLinSVC(HostFeatures(), svc_hparams={}, on_gpu=True).n_jobs == 1
