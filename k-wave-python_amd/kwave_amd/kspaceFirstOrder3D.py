"""Legacy 3-D entry point, python route (mirrors kwave/kspaceFirstOrder3D.py:133-211)."""
import warnings


def kspaceFirstOrder3D(kgrid, source, sensor, medium, simulation_options, execution_options, time_rev=False):
    """3-D k-space pseudospectral simulation; returns the sensor-data dict."""
    warnings.warn(
        "kspaceFirstOrder3D is deprecated. Use kspaceFirstOrder() from kwave.kspaceFirstOrder instead. See "
        "kwave.compat.options_to_kwargs() for migrating options.", FutureWarning, stacklevel=2)
    if kgrid.dim != 3:
        raise ValueError("kspaceFirstOrder3D needs a 3-D kgrid")
    if not execution_options.is_python_backend:
        raise NotImplementedError("only backend='python' (the native solver path, here on the GPU) is provided")
    from .solvers.native import run_python_backend

    return run_python_backend(kgrid, medium, source, sensor, simulation_options, execution_options)


def kspaceFirstOrder3DG(kgrid, source, sensor, medium, simulation_options, execution_options):
    """GPU alias (kwave/kspaceFirstOrder3D.py:23-76)."""
    execution_options.is_gpu_simulation = True
    return kspaceFirstOrder3D(kgrid, source, sensor, medium, simulation_options, execution_options)
