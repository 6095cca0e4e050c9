// Pass-kernel instances: 64-bit input and output (32- and 62-bit arithmetic).
#include "ntt_launch.hpp"
namespace nttk {
NTT_PASS_INSTANCES_MF(, uint64_t, uint64_t)
NTT_PASS_INSTANCES(, Mod64, uint64_t, uint64_t)
}
