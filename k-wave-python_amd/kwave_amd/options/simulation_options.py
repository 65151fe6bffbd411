"""Simulation options (the fields the native solver path reads).

Mirrors kwave/options/simulation_options.py:45-194: pml_size / pml_alpha,
use_sg, use_kspace, smooth_p0, pml_inside, data_cast.  HDF5 / C++ fields are
kept as plain attributes for call compatibility.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class SimulationOptions:
    simulation_type: object = None
    cart_interp: str = "linear"
    pml_inside: bool = True
    pml_alpha: float = 2.0
    save_to_disk: bool = False
    save_to_disk_exit: bool = False
    scale_source_terms: bool = True
    smooth_c0: bool = False
    smooth_rho0: bool = False
    smooth_p0: bool = True
    use_kspace: bool = True
    use_sg: bool = True
    use_fd: Optional[int] = None
    pml_auto: bool = False
    create_log: bool = False
    use_finite_difference: bool = False
    stream_to_disk: bool = False
    data_recast: Optional[bool] = False
    cartesian_interp: str = "linear"
    hdf_compression_level: Optional[int] = None
    data_cast: str = "off"
    pml_search_range: List[int] = field(default_factory=lambda: [10, 40])
    radial_symmetry: str = "WSWA-FFT"
    multi_axial_PML_ratio: float = 0.1
    data_path: Optional[str] = None
    input_filename: Optional[str] = None
    output_filename: Optional[str] = None
    pml_x_alpha: Optional[float] = None
    pml_y_alpha: Optional[float] = None
    pml_z_alpha: Optional[float] = None
    pml_size: Optional[List[int]] = None
    pml_x_size: Optional[int] = None
    pml_y_size: Optional[int] = None
    pml_z_size: Optional[int] = None

    def __post_init__(self):
        if self.cartesian_interp not in ("linear", "nearest"):
            raise ValueError("Optional input ''cartesian_interp'' must be set to ''linear'' or ''nearest''.")
        if self.data_cast not in ("off", "double", "single"):
            raise ValueError("Invalid input for ''data_cast''.")
        if self.data_cast == "double":
            self.data_cast = "off"
        if self.pml_size is not None:
            if isinstance(self.pml_size, int):
                self.pml_size = np.array([self.pml_size])
            if not isinstance(self.pml_size, (list, tuple, np.ndarray)):
                raise ValueError("Optional input ''PMLSize'' must be a integer array of 1, 2 or 3 dimensions.")
        self.pml_x_alpha = self.pml_alpha if self.pml_x_alpha is None else self.pml_x_alpha
        self.pml_y_alpha = self.pml_alpha if self.pml_y_alpha is None else self.pml_y_alpha
        self.pml_z_alpha = self.pml_alpha if self.pml_z_alpha is None else self.pml_z_alpha
