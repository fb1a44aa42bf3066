// flux_face_3d_e4.hip -- k_face_flux_batch<3, FC, false> (Eilmer 4 numerics).
#include "flux_face.hpp"

namespace gdtk {

hipError_t launch_face_flux_3d_e4(const BlockDev* tab, const FaceBatch& fb, int nblocks, int flux_calc,
                                  double eps, const IdealGasDev& gas, int64_t total, hipStream_t s) {
    return launch_face_flux_fc<3, false>(tab, fb, nblocks, flux_calc, eps, gas, total, s);
}

}  // namespace gdtk
