// Instantiation of the N = 32 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n32(int prec)
{
    if (prec) return new GridImpl<float, 32>();
    return new GridImpl<double, 32>();
}
} // namespace spm
