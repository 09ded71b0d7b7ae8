// Instantiation of the N = 96 passes and their tunable shapes (threads per
// pencil of the row, j and i passes; one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int shape)
{
    switch (shape) {
    case 1: return new GridImpl<T, 96, 4, 8, 8>();
    case 2: return new GridImpl<T, 96, 8, 8, 4>();
    case 3: return new GridImpl<T, 96, 4, 4, 4>();
    default: return new GridImpl<T, 96>();
    }
}
} // namespace

const int kShapes96[][3] = {{8, 8, 8}, {4, 8, 8}, {8, 8, 4}, {4, 4, 4}};

GridIface *make_grid_n96(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n96(const int (**s)[3])
{
    *s = kShapes96;
    return (int)(sizeof(kShapes96) / sizeof(kShapes96[0]));
}
} // namespace spm
