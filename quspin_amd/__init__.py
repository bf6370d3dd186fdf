"""MI355X-native implementation of QuSpin's symmetry-reduced H.v path.

``basis``      device bases (spin_basis_general, fermion bases, spin_basis_1d)
``operators``  hamiltonian (device CSR) and quantum_LinearOperator (fused
               matrix-free kernel)
``tools``      expm_multiply_parallel, lanczos_*, matvec

All compute runs in ``libquspin_hip.so`` (C-ABI in ``include/quspin_hip.h``).
"""

__version__ = "0.1.0"

from . import _lib  # noqa: F401


def library_path():
    return _lib.LIB_PATH
