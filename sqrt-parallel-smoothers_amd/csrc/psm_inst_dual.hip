// Tangent (forward-mode) instantiations, square-root form only:
// DualT<double> for (nx, ny) in {(1,1), (2,1), (2,2), (3,2), (4,2), (5,2)} and
// DualT<float> for (5, 2).  Split in four translation units (-DPSM_DUAL_PART=k)
// so that they compile in parallel.
#include "psm_impl.hpp"

#if PSM_DUAL_PART == 1
PSM_INSTANTIATE(psm_ops_d64s_1_1, psm::DualT<double>, 1, 1, true)
PSM_INSTANTIATE(psm_ops_d64s_2_1, psm::DualT<double>, 2, 1, true)
PSM_INSTANTIATE(psm_ops_d64s_2_2, psm::DualT<double>, 2, 2, true)
#elif PSM_DUAL_PART == 2
PSM_INSTANTIATE(psm_ops_d64s_3_2, psm::DualT<double>, 3, 2, true)
PSM_INSTANTIATE(psm_ops_d64s_4_2, psm::DualT<double>, 4, 2, true)
#elif PSM_DUAL_PART == 3
PSM_INSTANTIATE(psm_ops_d64s_5_2, psm::DualT<double>, 5, 2, true)
#elif PSM_DUAL_PART == 4
PSM_INSTANTIATE(psm_ops_d32s_5_2, psm::DualT<float>, 5, 2, true)
#endif
