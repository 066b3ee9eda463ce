"""Linearization methods (parsmooth/linearization/__init__.py:1-5)."""
from ._common import get_conditional_model
from ._methods import extended, cubature, gauss_hermite, unscented
