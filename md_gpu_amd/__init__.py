"""MI355X-native MD step engine (drop-in for the single-GPU step path of md_gpu).

Submodules:
  systems  seeded generators of the benchmark/test systems and the force-field record
  engine   ctypes wrapper of libmdgpu.so (C ABI in include/mdgpu/md_api.h)
  fileio   reader/writer of the reference's .inp / .fld / .cfg files
  fit      torch.autograd wrapper of the force-matching gradient
  build    hipcc build of libmdgpu.so and the athom tool
"""
from . import systems  # noqa: F401

__all__ = ["systems", "engine", "fileio", "fit", "build"]
