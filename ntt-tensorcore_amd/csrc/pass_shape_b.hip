// Shape-specialised pass kernels (32-bit coefficients), group B.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_B(NTT_SHAPE_INSTANCE)
}
