// Instantiation of the N = 128 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n128(int prec)
{
    if (prec) return new GridImpl<float, 128>();
    return new GridImpl<double, 128>();
}
} // namespace spm
