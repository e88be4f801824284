// qnd_kernels_0_double.hip -- one share of libqnd's line-pass kernels
// (QND_PASS_SET_0, double), compiled in a translation unit of its own so
// that the kernels build in parallel.
#include "qnd_device.hpp"

namespace qnd {
QND_PASS_SET_0(QND_PASS_INSTANTIATE, double)
QND_GROUP_SET(QND_GROUP_INSTANTIATE, double)
}  // namespace qnd
