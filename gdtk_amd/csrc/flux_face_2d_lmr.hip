// flux_face_2d_lmr.hip -- k_face_flux_batch<2, FC, true> (Eilmer 5 numerics).
#include "flux_face.hpp"

namespace gdtk {

hipError_t launch_face_flux_2d_lmr(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc,
                                   double eps, const IdealGasDev& gas, int64_t total, hipStream_t s) {
    return launch_face_flux_fc<2, true>(tab, fb, nblocks, flux_calc, eps, gas, total, s);
}

}  // namespace gdtk
