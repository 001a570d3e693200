"""speedynn_amd: greedy kNN wrapper feature selection on the AMD MI355X.

Drop-in for the selector of gieseke/speedynn (``speedynn.base``): the same
``IncrementalNNFeatureSelection`` API over a HIP engine (libspeedynn_amd.so,
C ABI in include/speedynn.h).
"""
__version__ = "1.0"

from .base import IncrementalNNFeatureSelection  # noqa: F401
