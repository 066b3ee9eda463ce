"""The reference's module path parsmooth/linearization/_gh.py:13-41: `linearize` is
parsmooth.linearization.gauss_hermite (see _methods)."""
from ._methods import gauss_hermite as linearize  # noqa: F401
