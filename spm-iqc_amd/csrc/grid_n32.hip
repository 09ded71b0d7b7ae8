// Instantiation of the N = 32 passes (one translation unit per N so they
// compile in parallel).
#include "grid_impl.h"

namespace spm {
namespace {
template <typename T> static GridIface *make(int) { return new GridImpl<T, 32>(); }
} // namespace

const int kShapes32[][3] = {{4, 4, 4}};

GridIface *make_grid_n32(int prec, int shape) { return prec ? make<float>(shape) : make<double>(shape); }

int grid_shapes_n32(const int (**s)[3])
{
    *s = kShapes32;
    return (int)(sizeof(kShapes32) / sizeof(kShapes32[0]));
}
} // namespace spm
