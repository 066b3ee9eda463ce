"""Linearization methods (parsmooth/linearization/__init__.py:1-5)."""
from ._methods import extended, cubature, gauss_hermite, unscented


def get_conditional_model(*_args, **_kwargs):
    """Conditional-moments models are not supported by the HIP path."""
    raise NotImplementedError("get_conditional_model / ConditionalMomentsModel are not supported "
                              "by the HIP implementation")
