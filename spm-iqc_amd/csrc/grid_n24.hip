// Instantiation of the N = 24 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 24>(); }
} // namespace

const int kShapes24[][3] = {{2, 2, 2}};

GridIface *make_grid_n24(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n24(const int (**s)[3])
{
    *s = kShapes24;
    return (int)(sizeof(kShapes24) / sizeof(kShapes24[0]));
}
} // namespace spm
