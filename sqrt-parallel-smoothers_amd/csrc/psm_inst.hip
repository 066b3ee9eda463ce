// One (dtype, form, nx) slice of the kernel instantiations; compiled several
// times with -DPSM_R=<float|double> -DPSM_SQRT=<0|1> -DPSM_NX=<n> -DPSM_TAG=<tag>.
#include "psm_impl.hpp"

#define PSM_CAT4_(a, b, c, d) a##b##_##c##_##d
#define PSM_CAT4(a, b, c, d) PSM_CAT4_(a, b, c, d)
#define PSM_FN(NY) PSM_CAT4(psm_ops_, PSM_TAG, PSM_NX, NY)

PSM_INSTANTIATE(PSM_FN(1), PSM_R, PSM_NX, 1, PSM_SQRT)
PSM_INSTANTIATE(PSM_FN(2), PSM_R, PSM_NX, 2, PSM_SQRT)
PSM_INSTANTIATE(PSM_FN(3), PSM_R, PSM_NX, 3, PSM_SQRT)
PSM_INSTANTIATE(PSM_FN(4), PSM_R, PSM_NX, 4, PSM_SQRT)
