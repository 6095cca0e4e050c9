// Pipelined pass kernels for tiles of 2^14 words (two workgroups per CU).
#include "ntt_launch.hpp"
namespace nttk {
NTT_PIPE_SHAPES_B(NTT_PIPE_INSTANCE)
}
