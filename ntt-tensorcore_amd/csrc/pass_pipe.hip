// Pipelined pass kernels: several tiles per workgroup, the next tile's loads
// in flight during the current tile's rounds (32-bit coefficients); tiles of
// 2^12 and 2^13 words.
#include "ntt_launch.hpp"
namespace nttk {
NTT_PIPE_SHAPES_A(NTT_PIPE_INSTANCE)
}
