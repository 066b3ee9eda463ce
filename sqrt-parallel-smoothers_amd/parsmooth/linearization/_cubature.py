"""The reference's module path parsmooth/linearization/_cubature.py:10-34: `linearize` is
parsmooth.linearization.cubature (see _methods)."""
from ._methods import cubature as linearize  # noqa: F401
