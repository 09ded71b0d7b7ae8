// Instantiation of the N = 64 passes and their tunable shapes (threads per
// pencil of the row, j and i passes; one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int shape)
{
    switch (shape) {
    case 1: return new GridImpl<T, 64, 8, 4, 4>();
    case 2: return new GridImpl<T, 64, 4, 4, 8>();
    case 3: return new GridImpl<T, 64, 8, 8, 8>();
    default: return new GridImpl<T, 64>();
    }
}
} // namespace

const int kShapes64[][3] = {{4, 4, 4}, {8, 4, 4}, {4, 4, 8}, {8, 8, 8}};

GridIface *make_grid_n64(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n64(const int (**s)[3])
{
    *s = kShapes64;
    return (int)(sizeof(kShapes64) / sizeof(kShapes64[0]));
}
} // namespace spm
