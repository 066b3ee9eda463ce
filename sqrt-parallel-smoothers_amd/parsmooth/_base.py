"""Container types of the public API (parsmooth/_base.py:5-30 of the reference).

Arrays are torch tensors (device arrays on the HIP device); numpy arrays are
accepted wherever the reference accepts jax arrays and are moved to the device.
"""
from typing import Any, Callable, NamedTuple, Union


class MVNStandard(NamedTuple):
    mean: Any
    cov: Any


class MVNSqrt(NamedTuple):
    mean: Any
    chol: Any


class FunctionalModel(NamedTuple):
    function: Callable
    mvn: Union[MVNSqrt, MVNStandard]


class ConditionalMomentsModel(NamedTuple):
    conditional_mean: Callable
    conditional_covariance_or_cholesky: Callable


def are_inputs_compatible(*y):
    """Raise TypeError unless some consecutive pair of arguments shares its type
    (the reference's rule, parsmooth/_base.py:25-30); with fewer than two
    arguments there is no such pair, so it raises too."""
    types = [type(v) for v in y]
    if not any(a == b for a, b in zip(types[:-1], types[1:])):
        raise TypeError(f"All inputs should have the same type. {y} was given")
