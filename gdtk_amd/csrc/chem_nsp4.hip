// chem_nsp4.hip -- chemistry kernels instantiated for 4-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<4>(const ChemCall& c);
}  // namespace gdtk
