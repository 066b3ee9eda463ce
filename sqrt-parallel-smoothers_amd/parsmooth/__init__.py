"""parsmooth on MI355X: parallel square-root Gaussian filtering, smoothing and
iterated smoothing (IEKS / IPLS) with the public API of the reference library
(EEA-sensors/sqrt-parallel-smoothers), computed by hand-written HIP kernels
(libparsmooth_hip.so) on one GPU."""
from ._base import MVNStandard, MVNSqrt, FunctionalModel, ConditionalMomentsModel
from .methods import (filtering, smoothing, filter_smoother, iterated_smoothing, sampling, FixedIterations,
                      clear_caches)
from ._engine import set_fp32_fallback

__version__ = "1.0.0"
