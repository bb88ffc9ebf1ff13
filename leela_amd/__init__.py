"""MI355X-native backend for Leela's batched CNN leaf evaluation.

The C ABI (include/leela_amd.h, libleela_amd.so) replaces the reference's
OpenCL backend; ``backend.Backend`` is its Python binding and ``nets`` holds
the reference layer tables with synthetic He-normal weights.
"""
from .backend import (Backend, LeelaError, NET_POLICY, NET_VALUE, PREC_BF16,
                      PREC_FP16, PREC_FP8, PREC_FP32, PRECISIONS,
                      load_library)
from . import nets

__all__ = ["Backend", "LeelaError", "NET_POLICY", "NET_VALUE", "PREC_FP32",
           "PREC_BF16", "PREC_FP8", "PREC_FP16", "PRECISIONS", "load_library",
           "nets"]
