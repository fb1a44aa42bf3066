// chem_nsp8.hip -- chemistry kernels instantiated for 8-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<8>(const ChemCall& c);
}  // namespace gdtk
