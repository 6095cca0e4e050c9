// Work-queue kernels: one persistent launch for every pass of a batch
// (32-bit coefficients and arithmetic; matrix-core or VALU rounds).
#include "ntt_launch.hpp"
namespace nttk {
NTT_QUEUE_INSTANCES()
}
