"""Python shims over the TorchScript classes, with the names and signatures of the
reference's ``hastypy.ffi`` modules (python/python/hastypy/ffi/*.py)."""
