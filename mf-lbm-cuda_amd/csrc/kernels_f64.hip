// FP64 instantiation of the step kernels.  Built with -ffp-contract=off: each operation
// rounds once in source order, which makes the results bit-identical to the FP64 oracle.
#include "launch.inl"
namespace mflbm {
template struct Launch<double>;
}
