// Instantiation of the N = 128 passes and their tunable shapes (threads per
// pencil of the row, j and i passes; one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int shape)
{
    switch (shape) {
    case 1: return new GridImpl<T, 128, 16, 8, 8>();
    case 2: return new GridImpl<T, 128, 8, 8, 16>();
    case 3: return new GridImpl<T, 128, 16, 16, 16>();
    default: return new GridImpl<T, 128>();
    }
}
} // namespace

const int kShapes128[][3] = {{8, 8, 8}, {16, 8, 8}, {8, 8, 16}, {16, 16, 16}};

GridIface *make_grid_n128(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n128(const int (**s)[3])
{
    *s = kShapes128;
    return (int)(sizeof(kShapes128) / sizeof(kShapes128[0]));
}
} // namespace spm
