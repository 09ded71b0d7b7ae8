// Instantiation of the N = 16 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 16>(); }
} // namespace

const int kShapes16[][3] = {{4, 4, 4}};

GridIface *make_grid_n16(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n16(const int (**s)[3])
{
    *s = kShapes16;
    return (int)(sizeof(kShapes16) / sizeof(kShapes16[0]));
}
} // namespace spm
