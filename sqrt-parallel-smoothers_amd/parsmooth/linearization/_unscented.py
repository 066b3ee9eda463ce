"""The reference's module path parsmooth/linearization/_unscented.py:10-39: `linearize` is
parsmooth.linearization.unscented (see _methods)."""
from ._methods import unscented as linearize  # noqa: F401
