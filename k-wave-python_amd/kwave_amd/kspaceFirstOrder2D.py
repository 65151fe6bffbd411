"""Legacy 2-D entry point, python route (mirrors kwave/kspaceFirstOrder2D.py:141-218).

Only the native-solver route exists here: ``execution_options.backend`` must be
"python" (the C++ binary backend is a different product and is not provided).
"""
import warnings


def kspaceFirstOrder2D(kgrid, source, sensor, medium, simulation_options, execution_options):
    """2-D k-space pseudospectral simulation; returns the sensor-data dict."""
    warnings.warn(
        "kspaceFirstOrder2D is deprecated. Use kspaceFirstOrder() from kwave.kspaceFirstOrder instead. See "
        "kwave.compat.options_to_kwargs() for migrating options.", FutureWarning, stacklevel=2)
    if kgrid.dim != 2:
        raise ValueError("kspaceFirstOrder2D needs a 2-D kgrid")
    if not execution_options.is_python_backend:
        raise NotImplementedError("only backend='python' (the native solver path, here on the GPU) is provided")
    from .solvers.native import run_python_backend

    return run_python_backend(kgrid, medium, source, sensor, simulation_options, execution_options)


def kspace_first_order_2d_gpu(kgrid, source, sensor, medium, simulation_options, execution_options):
    """GPU alias (kwave/kspaceFirstOrder2D.py:24-79)."""
    execution_options.is_gpu_simulation = True
    return kspaceFirstOrder2D(kgrid, source, sensor, medium, simulation_options, execution_options)
