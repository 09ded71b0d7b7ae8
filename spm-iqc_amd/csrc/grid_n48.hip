// Instantiation of the N = 48 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 48>(); }
} // namespace

const int kShapes48[][3] = {{4, 4, 4}};

GridIface *make_grid_n48(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n48(const int (**s)[3])
{
    *s = kShapes48;
    return (int)(sizeof(kShapes48) / sizeof(kShapes48[0]));
}
} // namespace spm
