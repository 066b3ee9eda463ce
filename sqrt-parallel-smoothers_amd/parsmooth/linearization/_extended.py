"""The reference's module path parsmooth/linearization/_extended.py:9-70: `linearize` is
parsmooth.linearization.extended (see _methods)."""
from ._methods import extended as linearize  # noqa: F401
