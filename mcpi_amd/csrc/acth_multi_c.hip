// Instantiations of the multi-vector H.v kernel (acth_multi.h) for M in {11 12 13}.
#include "acth_multi.h"

namespace mcpi {

mcpi_status launch_multi_c(const BasisDev &b, const HamCoefDev &c0, int rank_doubles, int nbatch, size_t cstride,
                             const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st)
{
    switch (b.M) {
    case 11: return launch_multi_m<11>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 12: return launch_multi_m<12>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 13: return launch_multi_m<13>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
