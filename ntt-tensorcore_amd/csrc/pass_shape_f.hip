// Shape-specialised pass kernels (32-bit coefficients), group F.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_F(NTT_SHAPE_INSTANCE)
}
