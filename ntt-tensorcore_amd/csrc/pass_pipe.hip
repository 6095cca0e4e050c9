// Pipelined pass kernels: several tiles per workgroup, the next tile's loads
// in flight during the current tile's rounds (32-bit coefficients).
#include "ntt_launch.hpp"
namespace nttk {
NTT_PIPE_SHAPES(NTT_PIPE_INSTANCE)
}
