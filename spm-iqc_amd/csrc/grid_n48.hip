// Instantiation of the N = 48 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n48(int prec)
{
    if (prec) return new GridImpl<float, 48>();
    return new GridImpl<double, 48>();
}
} // namespace spm
