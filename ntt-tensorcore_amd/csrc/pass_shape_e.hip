// Shape-specialised pass kernels (32-bit coefficients), group E.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_E(NTT_SHAPE_INSTANCE)
}
