// FP64 instantiation of the engine kernels.  Built with -ffp-contract=off:
// every multiply and add rounds separately, as in the reference Fortran,
// which keeps the drop-in path within rounding of the FP64 oracle.
#include "uasd_kernels_impl.hpp"
namespace uasd {
UASD_INSTANTIATE(double)
}
