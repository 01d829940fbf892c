"""The Python reconstruction layer of the reference (python/python/hastypy/base/:
opalg, solvers, proximal, sense, recon) on GPU-resident tensors: the operators call the
TorchScript classes of this package, so iterates, coil maps and Toeplitz kernels stay
in HBM across iterations."""
