// FP32 instantiation of the step kernels (populations, collision and recolouring in
// float; monitor reductions accumulate in double).  Contraction into FMA is allowed.
#include "launch.inl"
namespace mflbm {
template struct Launch<float>;
}
