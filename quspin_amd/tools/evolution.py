"""``quspin.tools.evolution.expm_multiply_parallel`` (tools/evolution.py:11-16)."""

from .expm_multiply_parallel_core import expm_multiply_parallel, ExpmMultiplyParallel  # noqa: F401

__all__ = ["expm_multiply_parallel", "ExpmMultiplyParallel"]
