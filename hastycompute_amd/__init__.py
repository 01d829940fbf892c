"""MI355X-native batched SENSE hot path of HastyCompute (Toeplitz SENSE normal operator,
ES-kernel NUFFT type-1/type-2, LLR singular value thresholding, device CG)."""

__version__ = "0.1.0"
