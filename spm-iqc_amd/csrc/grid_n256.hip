// Instantiation of the N = 256 passes, with the threads-per-pencil variants
// the tuner chooses from (one translation unit per N so they compile in parallel).
#include <cstdlib>

#include "grid_impl.h"

namespace spm {
GridIface *make_grid_n256(int prec)
{
    int v = 0;
    if (const char *e = std::getenv("SPM_VARIANT_256")) v = std::atoi(e);
    if (prec) {
        switch (v) {
        case 1: return new GridImpl<float, 256, 32, 16, 32>();
        case 2: return new GridImpl<float, 256, 16, 16, 32>();
        case 3: return new GridImpl<float, 256, 32, 16, 16>();
        default: return new GridImpl<float, 256>();
        }
    }
    switch (v) {
    case 1: return new GridImpl<double, 256, 32, 16, 32>();
    case 2: return new GridImpl<double, 256, 16, 16, 32>();
    case 3: return new GridImpl<double, 256, 32, 16, 16>();
    default: return new GridImpl<double, 256>();
    }
}
} // namespace spm
