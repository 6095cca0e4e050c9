// Pass-kernel instances: 32-bit arithmetic, 32-bit input and output.
#include "ntt_launch.hpp"
namespace nttk {
NTT_PASS_INSTANCES_MF(, uint32_t, uint32_t)
}
