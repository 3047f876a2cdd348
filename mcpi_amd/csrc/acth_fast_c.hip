// Instantiations of the specialised H.v kernels (acth_fast.h) for M in {11 12 13}.
#include "acth_fast.h"

namespace mcpi {

mcpi_status launch_fast_c(const BasisDev &b, int maps, const HamCoefDev &c, int skip, int rank_doubles,
                             const cplx *C, cplx *out, hipStream_t st, int nbatch, size_t cstride,
                             int64_t vstride)
{
    switch (b.M) {
    case 11: return launch_fast_m<11>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 12: return launch_fast_m<12>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 13: return launch_fast_m<13>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
