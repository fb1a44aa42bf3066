// flux_species_lmr.hip -- the multi-species face kernels with Eilmer 5 (lmr) numerics, in a
// unit of their own so that the two numerics compile in parallel.
#include <hip/hip_runtime.h>

#include "flux_species.hpp"
#include "launchers.hpp"

namespace gdtk {

hipError_t launch_face_flux_species_lmr(const BlockDev* tab, const FaceBatch& fb, int nblocks, int dims, int flux_calc,
                                       double eps, int nsp, const ChemDev* cd, int64_t total, hipStream_t s) {
    if (nsp == 2) face_species<2, true>(tab, fb, nblocks, total, dims, flux_calc, eps, cd, s);
    else if (nsp == 5) face_species<5, true>(tab, fb, nblocks, total, dims, flux_calc, eps, cd, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace gdtk
