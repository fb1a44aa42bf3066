// chem_nsp7.hip -- chemistry kernels instantiated for 7-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<7>(const ChemCall& c);
}  // namespace gdtk
