// qnd_kernels_0_float.hip -- one share of libqnd's line-pass kernels
// (QND_PASS_SET_0, float), compiled in a translation unit of its own so
// that the kernels build in parallel.
#include "qnd_device.hpp"

namespace qnd {
QND_PASS_SET_0(QND_PASS_INSTANTIATE, float)
QND_GROUP_SET(QND_GROUP_INSTANTIATE, float)
}  // namespace qnd
