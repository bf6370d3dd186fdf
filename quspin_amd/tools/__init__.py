"""Tools on the H.v path: ``evolution.expm_multiply_parallel``, ``lanczos``,
``misc.matvec``."""

from . import evolution, lanczos, misc  # noqa: F401
from .misc import matvec, get_matvec_function  # noqa: F401
