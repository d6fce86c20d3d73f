"""Physics helpers: attenuation maps."""

from .attenuation import AttenuationVolume, DetectorXRF  # noqa: F401
