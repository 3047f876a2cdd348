// Instantiations of the multi-vector H.v kernel (acth_multi.h) for M in {7 8 9 10}.
#include "acth_multi.h"

namespace mcpi {

mcpi_status launch_multi_b(const BasisDev &b, const HamCoefDev &c0, int rank_doubles, int nbatch, size_t cstride,
                             const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st)
{
    switch (b.M) {
    case 7: return launch_multi_m<7>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 8: return launch_multi_m<8>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 9: return launch_multi_m<9>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 10: return launch_multi_m<10>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
