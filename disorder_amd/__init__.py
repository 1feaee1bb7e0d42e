"""MI355X-native DISORDER aligned-reconstruction hot path (gfx950 HIP kernels
behind the C-ABI of include/disorder.h)."""
from .ops import (CGsolver, CGsolver_batched, autotune, coil_apply, coil_precond, gather_samples, resampling, coil_basis, coil_gram, compressCoils, Context, Encoding, LMsolver, colmajor, computeEnergy, decode, default_context,
                  empty_colmajor, encode, encodedecode, is_colmajor, lm_partials, sincRigidTransform,
                  sincRigidTransformGradient)

__all__ = ["CGsolver_batched", "autotune", "coil_precond", "gather_samples", "resampling", "coil_apply", "coil_basis", "coil_gram", "compressCoils", "LMsolver", "lm_partials", "sincRigidTransformGradient", "CGsolver", "Context", "Encoding", "colmajor", "computeEnergy", "decode", "default_context", "empty_colmajor",
           "encode", "encodedecode", "is_colmajor", "sincRigidTransform"]
