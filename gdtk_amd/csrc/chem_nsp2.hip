// chem_nsp2.hip -- chemistry kernels instantiated for 2-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<2>(const ChemCall& c);
}  // namespace gdtk
