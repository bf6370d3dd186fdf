"""``quspin.tools.evolution.expm_multiply_parallel`` (tools/evolution.py:11-16)."""

from .expm_multiply_parallel_core import (  # noqa: F401
    ExpmMultiplyParallel, expm_multiply_parallel, expm_multiply_sectors)

__all__ = ["expm_multiply_parallel", "ExpmMultiplyParallel", "expm_multiply_sectors"]
