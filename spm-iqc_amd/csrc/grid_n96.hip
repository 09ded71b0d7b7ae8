// Instantiation of the N = 96 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n96(int prec)
{
    if (prec) return new GridImpl<float, 96>();
    return new GridImpl<double, 96>();
}
} // namespace spm
