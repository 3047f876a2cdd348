// Instantiations of the specialised H.v kernels (acth_fast.h) for M in {7 8 9 10}.
#include "acth_fast.h"

namespace mcpi {

mcpi_status launch_fast_b(const BasisDev &b, int maps, const HamCoefDev &c, int skip, int rank_doubles,
                             const cplx *C, cplx *out, hipStream_t st, int nbatch, size_t cstride,
                             int64_t vstride)
{
    switch (b.M) {
    case 7: return launch_fast_m<7>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 8: return launch_fast_m<8>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 9: return launch_fast_m<9>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    case 10: return launch_fast_m<10>(b, maps, c, skip, rank_doubles, C, out, st, nbatch, cstride, vstride);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
