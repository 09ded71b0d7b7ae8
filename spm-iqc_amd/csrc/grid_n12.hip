// Instantiation of the N = 12 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 12>(); }
} // namespace

const int kShapes12[][3] = {{2, 2, 2}};

GridIface *make_grid_n12(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n12(const int (**s)[3])
{
    *s = kShapes12;
    return (int)(sizeof(kShapes12) / sizeof(kShapes12[0]));
}
} // namespace spm
