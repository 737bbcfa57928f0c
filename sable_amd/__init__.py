"""sable_amd — MI355X-native VBR SpMV/SpMM executor behind SABLE's Python API.

Host modules (no GPU needed): ``vbr``, ``fileio``, ``convert``,
``autopartition``, ``synth``.  GPU modules (need ``libsable_amd.so``, built by
``python -m sable_amd.build``; they raise ImportError without it): ``runtime``,
``codegen``, ``router``.
"""
from .vbr import VBR, validate_vbr  # noqa: F401

__version__ = "0.1.0"
