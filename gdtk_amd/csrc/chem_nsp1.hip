// chem_nsp1.hip -- chemistry kernels instantiated for 1-species mechanisms.
#include "chem_kernels.hpp"

namespace gdtk {
template hipError_t chem_run<1>(const ChemCall& c);
}  // namespace gdtk
