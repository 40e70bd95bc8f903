// Stand-in: the reference's PSD projection choices (utils/eigen_ext.hpp).
#pragma once

#include <Eigen/Core>

namespace ipc {
enum class PSDProjectionMethod { NONE, CLAMP, ABS };
} // namespace ipc
