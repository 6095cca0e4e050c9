// Shape-specialised matrix-core pass kernels for uint64 data (first and last
// passes), group B.
#include "ntt_launch.hpp"
namespace nttk {
NTT_SHAPES_U64_B(NTT_SHAPE_U64_INSTANCE)
}
