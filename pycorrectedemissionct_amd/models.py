"""Volume geometry (the part of corrct's ``models.py`` the projectors touch).

``VolumeGeometry`` mirrors ``models.py:421-551``: shape in XYZ order, ``shape_zxy``
(Z first, then X/Y swapped as corrct's projectors expect), ``is_square``,
``is_3D``, ``vox_size``. Cone/parallel ``ProjectionGeometry`` is ASTRA-only in
the reference and has no counterpart here.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np


@dataclass
class VolumeGeometry:
    """Store the volume geometry (``models.py:421-551``)."""

    _vol_shape_xyz: np.ndarray = field(default_factory=lambda: np.array([0, 0]))
    vox_size: float = 1.0

    def __post_init__(self):
        self._vol_shape_xyz = np.array(self._vol_shape_xyz, ndmin=1)

    def is_square(self) -> bool:
        return bool(self._vol_shape_xyz[0] == self._vol_shape_xyz[1])

    @property
    def shape_xyz(self) -> np.ndarray:
        return self._vol_shape_xyz

    @property
    def shape_zxy(self) -> np.ndarray:
        vol_shape_zyx = np.flip(self._vol_shape_xyz)
        return np.array([*vol_shape_zyx[:-2], vol_shape_zyx[-1], vol_shape_zyx[-2]], dtype=int)

    @property
    def mask_shape(self) -> np.ndarray:
        return self.shape_xyz[:2]

    @property
    def extent(self):
        half = self._vol_shape_xyz * self.vox_size / 2
        return [hs * sign for hs in half for sign in (-1, +1)]

    def is_3D(self) -> bool:
        return len(self._vol_shape_xyz) == 3 and self._vol_shape_xyz[-1] > 1

    def get_3d(self) -> "VolumeGeometry":
        if len(self._vol_shape_xyz) == 2:
            return replace(self, _vol_shape_xyz=np.concatenate((self._vol_shape_xyz, [1])))
        return replace(self)

    @staticmethod
    def get_default_from_data(data, data_format: str = "dvwu") -> "VolumeGeometry":
        return get_vol_geom_from_data(data, data_format=data_format)

    @staticmethod
    def get_default_from_volume(volume) -> "VolumeGeometry":
        return get_vol_geom_from_volume(volume)


def get_vol_geom_from_data(data, padding_u: int = 0, data_format: str = "dvwu", super_sampling: int = 1) -> VolumeGeometry:
    """Default geometry from a sinogram shape (``models.py:731-779``): X = Y = U (minus padding), Z = V."""
    shape = tuple(np.shape(data))
    fmt = data_format[-len(shape):]
    dims = {k: shape[i] for i, k in enumerate(fmt)}
    if "u" not in dims:
        raise ValueError(f"Could not determine data dimensions. Data shape: {shape}, data format: {data_format}")
    u = dims["u"] - 2 * int(padding_u)
    dims_xyz = [u, u]
    if "v" in dims:
        dims_xyz.append(dims["v"])
    return VolumeGeometry(_vol_shape_xyz=np.array(dims_xyz) * super_sampling, vox_size=1.0 / super_sampling)


def get_vol_geom_from_volume(volume) -> VolumeGeometry:
    """Default geometry from a volume array ``[..., rows, cols]`` (``models.py:782-793``)."""
    shape = np.shape(volume)
    if len(shape) < 2:
        raise ValueError(f"The volume should be at least 2-dimensional, but the following shape was passed: {shape}")
    return VolumeGeometry(np.array([shape[-2], shape[-1], *np.flip(shape[:-2])]))
