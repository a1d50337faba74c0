"""MI355X-native pixel-space covariance path of CosmoForge.

``Fisher``, ``Spectra`` and ``PICSLike`` keep the reference's constructors
(in-memory inputs) and getters; the work runs in libcfmi (HIP, gfx950)
through its C ABI.
"""

from .fisher import Fisher
from .picslike import LikelihoodResult, ParameterGrid, PICSLike
from .spectra import Spectra

__all__ = ["Fisher", "Spectra", "PICSLike", "ParameterGrid", "LikelihoodResult"]
