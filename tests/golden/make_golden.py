"""Writes tests/golden/*.npz from the FP64 oracle (oracle/sigkern.py).

The reference ships no golden vectors and cannot run in this environment
(CuPy + numba.cuda only), so the fixtures are produced by the oracle, whose
every function is pinned in tests/test_oracle.py by closed forms, brute force
and invariants.  Inputs are seeded (numpy default_rng) and stored with the
outputs.  Run: python tests/golden/make_golden.py
"""

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import sigkern as O  # noqa: E402


def c0():
  """Smallest benchmark configuration: SignatureKernel(n_levels=5, order=1,
  RBFKernel(1.0)); X = 10 x 50 x 5 and Y = 8 x 20 x 5 i.i.d. N(0,1) float64."""
  rng = np.random.default_rng(20240)
  X = rng.standard_normal((10, 50, 5))
  Y = rng.standard_normal((8, 20, 5))
  kw = dict(n_levels=5, order=1, kind='rbf', param=1.0)
  out = dict(
      X=X, Y=Y,
      K_XX=O.signature_kernel(X, **kw),
      K_XY=O.signature_kernel(X, Y, **kw),
      K_X_diag=O.signature_kernel(X, diag=True, **kw),
      K_XX_unnorm=O.signature_kernel(X, normalize=False, **kw),
      K_XY_unnorm=O.signature_kernel(X, Y, normalize=False, **kw),
      K_X_diag_unnorm=O.signature_kernel(X, diag=True, normalize=False, **kw),
      levels_XY=O.signature_kern(O.pairwise_gram(X, Y, 'rbf', 1.0), 5, 1, True, True),
      pde_XY_unnorm=O.signature_pde_kernel(X[:, :20] * 0.3, Y * 0.3, normalize=False),
  )
  np.savez_compressed(os.path.join(HERE, 'c0_signature_rbf.npz'), **out)


def higher_order():
  """Order-2 and full-order levels on short RBF / linear sequence pairs."""
  rng = np.random.default_rng(20241)
  X = rng.standard_normal((3, 9, 2)) * 0.6
  Y = rng.standard_normal((2, 7, 2)) * 0.6
  out = dict(X=X, Y=Y)
  for order in (2, 3):
    out[f'levels_rbf_o{order}'] = O.signature_kern(
        O.pairwise_gram(X, Y, 'rbf', 0.9), 4, order, True, True)
  out['levels_lin_full'] = O.signature_kern(
      O.pairwise_gram(X, Y, 'linear', 1.0), 4, -1, True, True)
  np.savez_compressed(os.path.join(HERE, 'higher_order.npz'), **out)


def pde_dyadic():
  """Signature-PDE values with dyadic refinement 0..2 (RBF, bandwidth 1)."""
  rng = np.random.default_rng(20242)
  X = O.brownian(rng, 3, 30, 4)
  Y = O.brownian(rng, 2, 25, 4)
  out = dict(X=X, Y=Y)
  for lam in range(3):
    out[f'pde_lam{lam}'] = O.signature_pde_kernel(X, Y, normalize=False, dyadic_order=lam)
  np.savez_compressed(os.path.join(HERE, 'pde_dyadic.npz'), **out)


if __name__ == '__main__':
  c0()
  higher_order()
  pde_dyadic()
  print('wrote', sorted(f for f in os.listdir(HERE) if f.endswith('.npz')))
