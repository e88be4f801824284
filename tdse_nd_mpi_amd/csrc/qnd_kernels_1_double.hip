// qnd_kernels_1_double.hip -- one share of libqnd's line-pass kernels
// (QND_PASS_SET_1, double), compiled in a translation unit of its own so
// that the kernels build in parallel.
#include "qnd_device.hpp"

namespace qnd {
QND_PASS_SET_1(QND_PASS_INSTANTIATE, double)
}  // namespace qnd
