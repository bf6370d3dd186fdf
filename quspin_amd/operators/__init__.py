"""Operators on the H.v path: ``hamiltonian`` (CSR in HBM) and
``quantum_LinearOperator`` (matrix-free fused kernel)."""

from .hamiltonian_core import hamiltonian, ishamiltonian  # noqa: F401
from .quantum_LinearOperator_core import quantum_LinearOperator, isquantum_LinearOperator  # noqa: F401
from ._devop import DeviceCSR, MatrixFreeOp  # noqa: F401

__all__ = ["hamiltonian", "ishamiltonian", "quantum_LinearOperator", "isquantum_LinearOperator", "DeviceCSR",
           "MatrixFreeOp"]
