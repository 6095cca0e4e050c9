// Shape-specialised pass kernels (32-bit coefficients), group C.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_C(NTT_SHAPE_INSTANCE)
}
