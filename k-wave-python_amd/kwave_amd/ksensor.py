"""Sensor container (mirrors kwave/ksensor.py:9-71, fields used by the solver)."""


class kSensor:
    def __init__(self, mask=None, record=None):
        self._mask = mask
        self.record = record
        self._record_start_index = 1
        self.directivity = None
        self.frequency_response = None

    @property
    def mask(self):
        return self._mask

    @mask.setter
    def mask(self, val):
        self._mask = val

    @property
    def record_start_index(self):
        return self._record_start_index

    @record_start_index.setter
    def record_start_index(self, val):
        self._record_start_index = int(round(val))
