"""Medium container (mirrors kwave/kmedium.py:22-60, fields used by the solver)."""
from dataclasses import dataclass
from typing import Optional, Union

import numpy as np

from .enums import AlphaMode


def _to_alpha_mode(value):
    if value is None or isinstance(value, AlphaMode):
        return value
    try:
        return AlphaMode(value)
    except (ValueError, TypeError):
        raise ValueError(
            f"medium.alpha_mode must be an AlphaMode enum value or one of 'no_absorption', 'no_dispersion', 'stokes', got {value!r}"
        ) from None


@dataclass
class kWaveMedium:
    sound_speed: Union[float, int, np.ndarray]
    sound_speed_ref: Optional[Union[float, int, np.ndarray]] = None
    density: Optional[Union[float, int, np.ndarray]] = None
    alpha_coeff: Optional[Union[float, int, np.ndarray]] = None
    alpha_power: Optional[Union[float, int, np.ndarray]] = None
    alpha_mode: Optional[Union[AlphaMode, str]] = None
    alpha_filter: Optional[np.ndarray] = None
    alpha_sign: Optional[np.ndarray] = None
    BonA: Optional[Union[float, int, np.ndarray]] = None
    absorbing: bool = False
    stokes: bool = False

    def __post_init__(self):
        self.sound_speed = np.atleast_1d(self.sound_speed)
        self.alpha_mode = _to_alpha_mode(self.alpha_mode)
