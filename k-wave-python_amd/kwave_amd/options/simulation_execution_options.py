"""Execution options (the fields the native solver path reads).

Mirrors kwave/options/simulation_execution_options.py:13-135: ``backend``
("python" routes kspaceFirstOrder2D/3D to the native solver, here the HIP
solver) and ``is_gpu_simulation``.  Binary/CLI options of the C++ backend are
kept as plain attributes for call compatibility and are not used.
"""
import os
from typing import Optional


class SimulationExecutionOptions:
    def __init__(self, is_gpu_simulation: bool = False, binary_path: Optional[str] = None, binary_dir: Optional[str] = None,
                 binary_name: Optional[str] = None, kwave_function_name: Optional[str] = "kspaceFirstOrder3D",
                 delete_data: bool = True, device_num: Optional[int] = None, num_threads: Optional[int] = None,
                 thread_binding: Optional[bool] = None, system_call: Optional[str] = None, verbose_level: int = 0,
                 auto_chunking: Optional[bool] = True, show_sim_log: bool = True, checkpoint_interval: Optional[int] = None,
                 checkpoint_timesteps: Optional[int] = None, checkpoint_file=None, backend: Optional[str] = None):
        self.backend = backend
        self.is_gpu_simulation = is_gpu_simulation
        self.binary_path = binary_path
        self.binary_dir = binary_dir
        self.binary_name = binary_name
        self.kwave_function_name = kwave_function_name
        self.delete_data = delete_data
        self.device_num = device_num
        self.num_threads = num_threads if num_threads is not None else os.cpu_count()
        self.thread_binding = thread_binding
        self.system_call = system_call
        self.verbose_level = verbose_level
        self.auto_chunking = auto_chunking
        self.show_sim_log = show_sim_log
        self.checkpoint_interval = checkpoint_interval
        self.checkpoint_timesteps = checkpoint_timesteps
        self.checkpoint_file = checkpoint_file
        if checkpoint_file is not None and checkpoint_interval is None and checkpoint_timesteps is None:
            raise ValueError("One of checkpoint_interval or checkpoint_timesteps must be set when checkpoint_file is set.")

    @property
    def backend(self):
        return self._backend

    @backend.setter
    def backend(self, value):
        valid = [None, "OMP", "CUDA", "python"]
        if value not in valid:
            raise ValueError(f"Backend must be one of {valid}, got '{value}'")
        self._backend = value

    @property
    def is_python_backend(self) -> bool:
        return self._backend == "python"

    @property
    def verbose_level(self):
        return self._verbose_level

    @verbose_level.setter
    def verbose_level(self, value):
        if not (isinstance(value, int) and 0 <= value <= 2):
            raise ValueError("Verbose level must be between 0 and 2")
        self._verbose_level = value
