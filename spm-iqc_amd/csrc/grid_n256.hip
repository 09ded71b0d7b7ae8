// Instantiation of the N = 256 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n256(int prec)
{
    if (prec) return new GridImpl<float, 256>();
    return new GridImpl<double, 256>();
}
} // namespace spm
