// Instantiation of the N = 16 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n16(int prec)
{
    if (prec) return new GridImpl<float, 16>();
    return new GridImpl<double, 16>();
}
} // namespace spm
