from . import kernels
