// Instantiation of the N = 4 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n4(int prec)
{
    if (prec) return new GridImpl<float, 4>();
    return new GridImpl<double, 4>();
}
} // namespace spm
