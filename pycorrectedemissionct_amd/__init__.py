"""MI355X-native attenuation-corrected XRF projector and solvers.

A drop-in projector backend for corrct's ``ProjectorUncorrected`` /
``ProjectorAttenuationXRF`` API and its ``SIRT`` / ``MLEM`` / ``PDHG`` solvers:
the hot path (attenuation maps, forward projection, back-projection and the
solver updates) runs in hand-written HIP kernels for gfx950, reached through
the C ABI of ``libcorrct_hip.so`` (``include/corrct_hip.h``); torch tensors
only carry device memory and the stream.
"""

__version__ = "0.1.0"

from . import batch, filters, models, operators, projectors, solvers, data_terms, regularizers, testing, warmstart  # noqa: F401
from ._projector_backends import ProjectorBackend, ProjectorBackendHIP, compute_attenuation  # noqa: F401
from .physics.attenuation import AttenuationVolume  # noqa: F401
from .projectors import ProjectorAttenuationXRF, ProjectorMatrix, ProjectorUncorrected  # noqa: F401
from .solvers import FBP, FISTA, MLEM, PDHG, SART, SIRT, SolutionInfo  # noqa: F401
