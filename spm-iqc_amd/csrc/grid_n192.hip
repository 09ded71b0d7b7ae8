// Instantiation of the N = 192 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 192>(); }
} // namespace

const int kShapes192[][3] = {{16, 16, 16}};

GridIface *make_grid_n192(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n192(const int (**s)[3])
{
    *s = kShapes192;
    return (int)(sizeof(kShapes192) / sizeof(kShapes192[0]));
}
} // namespace spm
