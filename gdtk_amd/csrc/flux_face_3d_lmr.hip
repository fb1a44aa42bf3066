// flux_face_3d_lmr.hip -- k_face_flux_batch<3, FC, true> (Eilmer 5 numerics).
#include "flux_face.hpp"

namespace gdtk {

hipError_t launch_face_flux_3d_lmr(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc,
                                   double eps, const IdealGasDev& gas, int64_t total, hipStream_t s) {
    return launch_face_flux_fc<3, true>(tab, fb, nblocks, flux_calc, eps, gas, total, s);
}

}  // namespace gdtk
