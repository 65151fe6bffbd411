"""kwave_amd: an MI355X-native k-space pseudospectral acoustic solver.

The public surface mirrors k-wave-python's native solver path
(``kspaceFirstOrder2D/3D``, ``kspaceFirstOrder``, ``Simulation``) so that a
caller can switch imports.  All time stepping runs in hand-written gfx950 HIP
kernels reached through the C ABI in ``libkwgpu.so``; there is no CPU solver in
this package.
"""
__version__ = "0.1.0"
