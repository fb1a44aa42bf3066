// chem_nsp5.hip -- chemistry kernels instantiated for 5-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<5>(const ChemCall& c);
}  // namespace gdtk
