"""MI355X-native RBF-FD shallow-water RK4 hot path.

Modules:
  host     C host (libswe_host.so): reader, RCM, cfdl/sfdl layouts, serial patch
  runtime  HIP C-ABI (libswe_hip.so): device contexts, drop-in RK substeps,
           fused graph-captured steps, ensembles
  serve    continuous-batching forecast server (libswe_serve.so)
  mesh     seeded generator of inputs in the reference's binary format
"""
from . import host, mesh, runtime, serve  # noqa: F401

__all__ = ["host", "mesh", "runtime", "serve"]
