"""``ksig``: the reference library's import name (tgcsaba/KSig,
ksig/__init__.py:1-6) for this MI355X build, so that code written against
the reference runs unchanged::

    import ksig
    from ksig.static.kernels import Kernel
    from ksig.models import PrecomputedFeatureLinSVC, PrecomputedKernelSVC
    from ksig.utils import _EPS, ArrayOnCPUOrGPU, ArrayOnGPU

Every module name of the reference resolves to its ``ksig_amd`` counterpart
(the same module objects, registered under both names): the signature
kernels compute on the GPU through libksig_hip.so.
"""

import importlib as _importlib
import sys as _sys

_MODULES = ("utils", "static", "static.kernels", "static.features", "algorithms", "kernels",
            "preprocessing", "projections", "models", "tune")

for _name in _MODULES:
  _sys.modules[__name__ + "." + _name] = _importlib.import_module("ksig_amd." + _name)

from ksig_amd import (algorithms, kernels, models, preprocessing, projections,  # noqa: E402
                      static, tune, utils)
from ksig_amd import __version__  # noqa: E402,F401
