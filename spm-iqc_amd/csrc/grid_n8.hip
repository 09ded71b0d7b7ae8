// Instantiation of the N = 8 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 8>(); }
} // namespace

const int kShapes8[][3] = {{2, 2, 2}};

GridIface *make_grid_n8(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n8(const int (**s)[3])
{
    *s = kShapes8;
    return (int)(sizeof(kShapes8) / sizeof(kShapes8[0]));
}
} // namespace spm
