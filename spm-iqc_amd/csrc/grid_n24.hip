// Instantiation of the N = 24 passes (one translation unit per N so they compile in parallel).
#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n24(int prec)
{
    if (prec) return new GridImpl<float, 24>();
    return new GridImpl<double, 24>();
}
} // namespace spm
