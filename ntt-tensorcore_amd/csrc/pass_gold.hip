// Pass-kernel instances: Goldilocks arithmetic (q = 2^64 - 2^32 + 1), 64-bit input and output.
#include "ntt_launch.hpp"
namespace nttk {
NTT_PASS_INSTANCES(, ModGold, uint64_t, uint64_t)
}
