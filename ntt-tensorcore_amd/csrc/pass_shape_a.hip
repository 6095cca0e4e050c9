// Shape-specialised pass kernels (32-bit coefficients), group A.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_A(NTT_SHAPE_INSTANCE)
}
