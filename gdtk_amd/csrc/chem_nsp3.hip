// chem_nsp3.hip -- chemistry kernels instantiated for 3-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<3>(const ChemCall& c);
}  // namespace gdtk
