// chem_nsp6.hip -- chemistry kernels instantiated for 6-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<6>(const ChemCall& c);
}  // namespace gdtk
