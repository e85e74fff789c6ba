"""ksig_amd — MI355X (gfx950) signature-kernel hot path with the ksig API.

Drop-in for ``ksig.kernels.SignatureKernel`` / ``SignaturePDEKernel`` with
``ksig.static.kernels.RBFKernel`` / ``LinearKernel`` and the functional
``ksig.algorithms.signature_kern`` / ``signature_kern_pde``::

    import ksig_amd as ksig
    K = ksig.kernels.SignatureKernel(5, static_kernel=ksig.static.kernels.RBFKernel())(X)

The package ``ksig`` next to this one registers these modules under the
reference's own import name (``import ksig``, ``from ksig.models import ...``).
"""

from . import utils
from . import static
from . import algorithms
from . import kernels
from . import tune
from . import models
from . import projections
from . import preprocessing

__version__ = "0.1.0"
