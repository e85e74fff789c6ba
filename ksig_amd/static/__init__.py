from . import kernels
from . import features
