"""Device bases mirroring ``quspin.basis`` for the H.v path."""

from .base_general import basis_general, hcb_basis_general, GeneralBasisWarning  # noqa: F401
from .spin import spin_basis_general  # noqa: F401
from .fermion import spinless_fermion_basis_general, spinful_fermion_basis_general  # noqa: F401
from .basis_1d import spin_basis_1d  # noqa: F401

__all__ = ["basis_general", "hcb_basis_general", "spin_basis_general", "spinless_fermion_basis_general",
           "spinful_fermion_basis_general", "spin_basis_1d", "GeneralBasisWarning"]
