// qnd_kernels_1_x.hip -- one share of libqnd's line-pass kernels
// (QND_PASS_SET_1) for complex64 states held in FP64 registers (k_pass_x),
// compiled in a translation unit of its own so that the kernels build in
// parallel.
#include "qnd_device.hpp"

namespace qnd {
QND_PASS_SET_1(QND_PASSX_INSTANTIATE, float)
}  // namespace qnd
