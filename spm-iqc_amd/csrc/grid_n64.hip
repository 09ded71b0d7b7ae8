// Instantiation of the N = 64 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n64(int prec)
{
    if (prec) return new GridImpl<float, 64>();
    return new GridImpl<double, 64>();
}
} // namespace spm
