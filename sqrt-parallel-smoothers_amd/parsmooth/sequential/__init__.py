"""Sequential filter and smoother (parsmooth/sequential of the reference) on
the HIP device, one lane per sequence."""
from ._filtering import filtering
from ._smoothing import smoothing
