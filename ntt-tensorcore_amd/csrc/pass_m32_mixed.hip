// Pass-kernel instances: 32-bit arithmetic, 32-bit input with 64-bit output and the reverse.
#include "ntt_launch.hpp"
namespace nttk {
NTT_PASS_INSTANCES_MF(, uint32_t, uint64_t)
NTT_PASS_INSTANCES_MF(, uint64_t, uint32_t)
}
