// FP32 instantiation of the engine kernels (gather and rotation in single
// precision; observables are reduced in FP64).  Contraction into FMA is
// allowed here.
#include "uasd_kernels_impl.hpp"
namespace uasd {
UASD_INSTANTIATE(float)
}
