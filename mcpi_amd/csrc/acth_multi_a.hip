// Instantiations of the multi-vector H.v kernel (acth_multi.h) for M in {3 4 5 6}.
#include "acth_multi.h"

namespace mcpi {

mcpi_status launch_multi_a(const BasisDev &b, const HamCoefDev &c0, int rank_doubles, int nbatch, size_t cstride,
                             const cplx *C0, cplx *out0, int64_t vstride, hipStream_t st)
{
    switch (b.M) {
    case 3: return launch_multi_m<3>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 4: return launch_multi_m<4>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 5: return launch_multi_m<5>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    case 6: return launch_multi_m<6>(b, c0, rank_doubles, nbatch, cstride, C0, out0, vstride, st);
    default: return MCPI_ERR_STATE;
    }
}

}  // namespace mcpi
