// Instantiation of the N = 8 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n8(int prec)
{
    if (prec) return new GridImpl<float, 8>();
    return new GridImpl<double, 8>();
}
} // namespace spm
