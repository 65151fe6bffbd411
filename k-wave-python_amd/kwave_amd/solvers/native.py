"""Legacy bridge kspaceFirstOrder2D/3D -> Simulation (mirrors kwave/solvers/native.py:6-30).

Deviation (documented): an integer ``SimulationOptions(pml_size=n)`` arrives as
a length-1 array; the reference then indexes axis 1 and fails
(kspace_solver.py:381); here a length-1 value is broadcast to every axis.
"""
from .kspace_solver import Simulation


def _per_axis(val, dim, default, cast):
    if val is None:
        return (default,) * dim
    if not hasattr(val, "__len__"):
        return (cast(val),) * dim
    vals = [cast(v) for v in val]
    if len(vals) == 1:
        return (vals[0],) * dim
    return tuple(vals)


def run_python_backend(kgrid, medium, source, sensor, simulation_options, execution_options):
    device = execution_options.device_num if getattr(execution_options, "device_num", None) is not None else "gpu"
    pml_size = _per_axis(simulation_options.pml_size, kgrid.dim, 20, int)
    pml_alpha = _per_axis(getattr(simulation_options, "pml_alpha", None), kgrid.dim, 2.0, float)
    return Simulation(
        kgrid, medium, source, sensor, device=device,
        use_sg=getattr(simulation_options, "use_sg", True),
        use_kspace=getattr(simulation_options, "use_kspace", True),
        smooth_p0=getattr(simulation_options, "smooth_p0", True),
        pml_size=pml_size, pml_alpha=pml_alpha,
    ).run()
