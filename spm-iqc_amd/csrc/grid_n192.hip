// Instantiation of the N = 192 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n192(int prec)
{
    if (prec) return new GridImpl<float, 192>();
    return new GridImpl<double, 192>();
}
} // namespace spm
