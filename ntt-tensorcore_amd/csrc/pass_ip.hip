// Pass kernels for first passes whose load forms a product or an inner
// product (32-bit storage, matrix-core rounds, four-step post).
#include "ntt_launch.hpp"
namespace nttk {
NTT_IP_SHAPES(NTT_IP_INSTANCE)
}
