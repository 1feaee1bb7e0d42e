"""MI355X-native DISORDER aligned-reconstruction hot path (gfx950 HIP kernels
behind the C-ABI of include/disorder.h)."""
from .ops import (CGsolver, Context, Encoding, colmajor, computeEnergy, decode, default_context, empty_colmajor,
                  encode, encodedecode, is_colmajor, sincRigidTransform)

__all__ = ["CGsolver", "Context", "Encoding", "colmajor", "computeEnergy", "decode", "default_context", "empty_colmajor",
           "encode", "encodedecode", "is_colmajor", "sincRigidTransform"]
