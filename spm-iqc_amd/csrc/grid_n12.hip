// Instantiation of the N = 12 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n12(int prec)
{
    if (prec) return new GridImpl<float, 12>();
    return new GridImpl<double, 12>();
}
} // namespace spm
