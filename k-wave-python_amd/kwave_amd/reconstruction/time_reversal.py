"""Time-reversal reconstruction (mirrors kwave/reconstruction/time_reversal.py:24-150).

The recorded pressure is flipped in time and re-imposed at the sensor points as
a dirichlet source; p_final of that run, times a compensation factor and
clipped at zero, is the reconstruction.  The forward and reverse runs both go
through the GPU solver.
"""
import numpy as np

from ..ksensor import kSensor
from ..ksource import kSource


class TimeReversal:
    def __init__(self, kgrid, medium, sensor, compensation_factor: float = 2.0):
        self.kgrid, self.medium, self.sensor = kgrid, medium, sensor
        self.compensation_factor = compensation_factor
        self._source = None
        self._new_sensor = None
        if sensor.mask is None:
            raise ValueError("Sensor mask must be set for time reversal. Use sensor.mask = ...")
        t_array = kgrid.t_array
        if t_array is None or isinstance(t_array, str):
            raise ValueError("t_array must be explicitly set for time reversal")
        if compensation_factor <= 0:
            raise ValueError("compensation_factor must be positive")
        if not np.any(sensor.mask):
            raise ValueError("Sensor mask must have at least one active point")
        if not np.array_equal(np.asarray(sensor.mask).shape, kgrid.N):
            raise ValueError(f"Sensor mask shape {np.asarray(sensor.mask).shape} does not match grid dimensions {kgrid.N}")
        self._passed_record = list(self.sensor.record) if self.sensor.record is not None else []
        if "p_final" not in self._passed_record:
            self._passed_record.append("p_final")

    def __call__(self, simulation_function, simulation_options, execution_options):
        if simulation_function is None:
            raise ValueError("simulation_function must be provided")
        if simulation_options is None:
            raise ValueError("simulation_options must be provided")
        if execution_options is None:
            raise ValueError("execution_options must be provided")
        if getattr(self.sensor, "recorded_pressure", None) is None:
            raise ValueError("Sensor must have recorded pressure data. Run a forward simulation first.")
        self._source = kSource()
        self._source.p_mask = self.sensor.mask
        self._source.p = np.flip(self.sensor.recorded_pressure, axis=1)
        self._source.p_mode = "dirichlet"
        self._new_sensor = kSensor(mask=self.sensor.mask, record=self._passed_record)
        result = simulation_function(self.kgrid, self._source, self._new_sensor, self.medium, simulation_options, execution_options)
        p0_recon = result["p_final"] if isinstance(result, dict) else result
        p0_recon = self.compensation_factor * p0_recon
        p0_recon[p0_recon < 0] = 0
        return p0_recon.T
