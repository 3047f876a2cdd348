// Instantiations of the specialised H.v kernels (acth_fast.h) for M in {3 4 5 6}.
#include "acth_fast.h"

namespace mcpi {

mcpi_status launch_fast_a(const BasisDev &b, int maps, const HamCoefDev &c, int skip, int rank_doubles,
                             const cplx *C, cplx *out, hipStream_t st, int nbatch, size_t cstride,
                             int64_t vstride)
{
    switch (b.M) {
    case 3: return launch_fast_m<3>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 4: return launch_fast_m<4>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 5: return launch_fast_m<5>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 6: return launch_fast_m<6>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
