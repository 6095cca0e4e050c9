// Shape-specialised pass kernels (32-bit coefficients), group D.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_D(NTT_SHAPE_INSTANCE)
}
