// Full density matrices from the representative terms, entry by entry.
//
// The density kernels sum representative terms only (compact layouts of
// rho.hip: class 0 r1d | r1o | R1 | R2 | R4 | R5 | R7, class 1 R3 | R6,
// class 2 the two-orbital pairs).  The reference fills the other entries of
// rho1 and rho2 by symmetry copies (onebodyMatrix.h:246-247;
// twobodyMatrix.h:1050, 1073-1076, 1113-1114, 1141-1144, 1171-1174,
// 1214-1217, 1335-1343, 1477-1480).  Those copies never write one entry
// twice: every index pattern of rho2[a + bM + cM^2 + dM^3] =
// <a*_a a*_b a_c a_d> belongs to exactly one rule,
//   a = b = c = d                              R1
//   {a, b} = {c, d}, a != b                     R2
//   a = b, c = d, a != c                        R3 (conjugate if a > c)
//   three equal indices x, one other y          R4 if x < y, R5 if x > y
//                                               (conjugate when the pair
//                                               {x, y} sits in the creators)
//   a = b, c != d, neither equal to a           R6 (conjugate with the
//                                               roles of creators and
//                                               annihilators exchanged)
//   a != b, c != d, one common index s          R7[x, y, s] (x the other
//                                               creator, y the other
//                                               annihilator; conjugate if
//                                               x > y)
//   a != b, c != d, no common index             two-orbital pairs
// so each entry can be computed from its own index, with no ordering
// between writes: rho2_entry below is that inverse map.
#pragma once

#include "mcpi_internal.h"

namespace mcpi {

// offsets of the class-0 representatives (P = M(M-1)/2)
struct RepLayout {
    int M, P;
    __host__ __device__ int r1o() const { return M; }
    __host__ __device__ int R1() const { return M + P; }
    __host__ __device__ int R2() const { return M + P + M; }
    __host__ __device__ int R4() const { return R2() + P; }
    __host__ __device__ int R5() const { return R4() + P; }
    __host__ __device__ int R7() const { return R5() + P; }
    __host__ __device__ int n_class0() const { return R7() + P * M; }
    __host__ __device__ int n_class1() const { return P + M * P; }
};

__host__ __device__ inline cplx conj_if(cplx z, bool c) { return c ? cmake(z.re, -z.im) : z; }

// rho1[k + M l] = <a*_k a_l>
__host__ __device__ inline cplx rho1_entry(int k, int l, const RepLayout &L, const cplx *A)
{
    if (k == l) return A[k];
    const bool c = k > l;
    const int lo = c ? l : k, hi = c ? k : l;
    return conj_if(A[L.r1o() + pair_index(lo, hi, L.M)], c);
}

// rho2[a + bM + cM^2 + dM^3] = <a*_a a*_b a_c a_d> from class 0 (A), class 1
// (OB: R3 | R6) and class 2 (TT: pair (k<s) x pair (q<l), P x P)
__host__ __device__ inline cplx rho2_entry(int a, int b, int c, int d, const RepLayout &L, const cplx *A,
                                           const cplx *OB, const cplx *TT)
{
    const int M = L.M, P = L.P;
    auto pr = [M](int x, int y) { return x < y ? pair_index(x, y, M) : pair_index(y, x, M); };
    if (a == b && c == d) {
        if (a == c) return A[L.R1() + a];                          // R1
        return conj_if(OB[pr(a, c)], a > c);                       // R3
    }
    if (a == b || c == d) {
        // one side holds a repeated index x; y, z are the two others
        const bool cre = a == b;             // creators repeated
        const int x = cre ? a : c;
        const int y = cre ? c : a, z = cre ? d : b;
        if (y == x || z == x) {
            // three equal indices x and one other o
            const int o = y == x ? z : y;
            if (x < o) return conj_if(A[L.R4() + pr(x, o)], !cre);  // R4
            return conj_if(A[L.R5() + pr(x, o)], cre);              // R5
        }
        return conj_if(OB[P + x * P + pr(y, z)], !cre);             // R6
    }
    // a != b and c != d
    const int common = (a == c) + (a == d) + (b == c) + (b == d);
    if (common == 2) return A[L.R2() + pr(a, b)];                   // R2
    if (common == 1) {
        const int s = (a == c || a == d) ? a : b;                   // the common index
        const int x = a == s ? b : a, y = c == s ? d : c;
        return conj_if(A[L.R7() + pr(x, y) * M + s], x > y);         // R7
    }
    return TT[(int64_t)pr(a, b) * P + pr(c, d)];                    // two-orbital pairs
}

}  // namespace mcpi
