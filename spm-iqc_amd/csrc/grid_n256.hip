// Instantiation of the N = 256 passes and their tunable shapes (threads per
// pencil of the row, j and i passes; one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int shape)
{
    switch (shape) {
    case 1: return new GridImpl<T, 256, 32, 16, 32>();
    case 2: return new GridImpl<T, 256, 16, 16, 32>();
    case 3: return new GridImpl<T, 256, 32, 16, 16>();
    case 4: return new GridImpl<T, 256, 64, 16, 16>();
    case 5: return new GridImpl<T, 256, 32, 32, 16>();
    case 6: return new GridImpl<T, 256, 32, 8, 16>();
    case 7: return new GridImpl<T, 256, 64, 32, 16>();
    case 8: return new GridImpl<T, 256, 16, 32, 16>();
    case 9: return new GridImpl<T, 256, 16, 8, 16>();
    default: return new GridImpl<T, 256>();
    }
}
} // namespace

const int kShapes256[][3] = {{16, 16, 16}, {32, 16, 32}, {16, 16, 32}, {32, 16, 16}, {64, 16, 16}, {32, 32, 16}, {32, 8, 16}, {64, 32, 16}, {16, 32, 16}, {16, 8, 16}};

GridIface *make_grid_n256(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n256(const int (**s)[3])
{
    *s = kShapes256;
    return (int)(sizeof(kShapes256) / sizeof(kShapes256[0]));
}
} // namespace spm
