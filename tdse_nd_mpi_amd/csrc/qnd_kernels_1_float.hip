// qnd_kernels_1_float.hip -- one share of libqnd's line-pass kernels
// (QND_PASS_SET_1, float), compiled in a translation unit of its own so
// that the kernels build in parallel.
#include "qnd_device.hpp"

namespace qnd {
QND_PASS_SET_1(QND_PASS_INSTANTIATE, float)
}  // namespace qnd
