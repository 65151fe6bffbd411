"""Source container (mirrors kwave/ksource.py:10-66, fields used by the solver)."""
import numpy as np


class kSource:
    def __init__(self):
        self._p0 = None
        self.p = None
        self.p_mask = None
        self.p_mode = None
        self.p_frequency_ref = None
        self.ux = None
        self.uy = None
        self.uz = None
        self.u_mask = None
        self.u_mode = None
        self.u_frequency_ref = None
        self.s_mask = None

    def is_p0_empty(self) -> bool:
        return self.p0 is None or len(self.p0) == 0 or (np.sum(self.p0 != 0) == 0)

    @property
    def p0(self):
        return self._p0

    @p0.setter
    def p0(self, val):
        if val is None or len(val) == 0:
            self._p0 = None
        else:
            self._p0 = val
