// Factorized pair form of a Hamiltonian (the pair-space H.v of schedule 7,
// acth_pair.hip), detected on the host from the pair matrix and the
// one-body table that build_coefficients folds from Ho and Hint.
//
// The form holds when h1 is diagonal and the pair matrix W is block
// diagonal over the connected components of its non-zero pattern with
// every block of rank one: W[P][Q] = za[P] zb[Q] for P, Q in one block.
// A contact interaction that conserves a pair charge has it (Lieb-Liniger's
// Hint = g delta(k + s, q + l) gives W = g/2 m m^T on each total momentum,
// m = 1 or 2).  Then sum_{P,Q} W a*_P a_Q = sum_r A_r^+ B_r with
// B_r = sum_{Q in r} zb[Q] a_Q.  Plain C++ (no HIP), so that the host
// tests compile it on its own.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace mcpi {

// zblk[P]: block of pair P (-1: no term); zord lists the pairs block by
// block, block r at zord[zoff[r] .. zoff[r+1]); za, zb the factors
template <class Z>
struct PairFactor {
    std::vector<Z> za, zb;
    std::vector<int32_t> zblk, zord, zoff;
};

// W(P, Q) = pw[Q * PWS + P] over NP pairs, h1[k][l] = ph1[l * HS + k] over
// M orbitals.  True (and f filled) when the factorized form holds, every
// entry of a block within 1e-13 of the block's largest entry; false (f
// empty) otherwise, or when W has no non-zero entry.  Each block's factors
// come from its largest entry W[P0][Q0]: za = W[:, Q0], zb = W[P0, :] /
// W[P0][Q0].
template <class Z>
bool pair_factorize(int M, int NP, int PWS, int HS, const Z *pw, const Z *ph1, PairFactor<Z> &f)
{
    f = PairFactor<Z>();
    auto nz = [](const Z &z) { return z.re != 0.0 || z.im != 0.0; };
    auto W = [&](int P, int Q) -> const Z & { return pw[(size_t)Q * PWS + P]; };
    auto make = [](double re, double im) {
        Z z;
        z.re = re;
        z.im = im;
        return z;
    };
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++)
            if (l != k && nz(ph1[(size_t)l * HS + k])) return false;
    // blocks: connected components of the non-zero pattern (union-find)
    std::vector<int> par(NP);
    for (int x = 0; x < NP; x++) par[x] = x;
    auto root = [&](int x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
    };
    std::vector<uint8_t> used(NP, 0);
    for (int P = 0; P < NP; P++)
        for (int Q = 0; Q < NP; Q++)
            if (nz(W(P, Q))) {
                used[P] = used[Q] = 1;
                par[root(P)] = root(Q);
            }
    std::vector<int32_t> blk(NP, -1);
    std::vector<int> rid(NP, -1);
    int R = 0;
    for (int P = 0; P < NP; P++)
        if (used[P]) {
            const int r = root(P);
            if (rid[r] < 0) rid[r] = R++;
            blk[P] = rid[r];
        }
    if (R == 0) return false;
    std::vector<Z> za(NP, make(0, 0)), zb(NP, make(0, 0));
    for (int r = 0; r < R; r++) {
        std::vector<int> mem;
        for (int P = 0; P < NP; P++)
            if (blk[P] == r) mem.push_back(P);
        int P0 = mem[0], Q0 = mem[0];
        double big = -1;
        for (int P : mem)
            for (int Q : mem) {
                const double a = std::hypot(W(P, Q).re, W(P, Q).im);
                if (a > big) big = a, P0 = P, Q0 = Q;
            }
        const Z w0 = W(P0, Q0);
        const double n0 = w0.re * w0.re + w0.im * w0.im;
        for (int P : mem) za[P] = W(P, Q0);
        for (int Q : mem) {   // W(P0, Q) / w0
            const Z w = W(P0, Q);
            zb[Q] = make((w.re * w0.re + w.im * w0.im) / n0, (w.im * w0.re - w.re * w0.im) / n0);
        }
        for (int P : mem)
            for (int Q : mem) {
                const Z w = W(P, Q);
                const double fr = za[P].re * zb[Q].re - za[P].im * zb[Q].im;
                const double fi = za[P].re * zb[Q].im + za[P].im * zb[Q].re;
                if (std::hypot(w.re - fr, w.im - fi) > 1e-13 * big) return false;
            }
    }
    std::vector<int32_t> off(R + 1, 0);
    for (int P = 0; P < NP; P++)
        if (blk[P] >= 0) off[blk[P] + 1]++;
    for (int r = 0; r < R; r++) off[r + 1] += off[r];
    std::vector<int32_t> ord(off[R], 0), fill(off.begin(), off.end() - 1);
    for (int P = 0; P < NP; P++)
        if (blk[P] >= 0) ord[fill[blk[P]]++] = P;
    f.za = std::move(za);
    f.zb = std::move(zb);
    f.zblk = std::move(blk);
    f.zord = std::move(ord);
    f.zoff = std::move(off);
    return true;
}

}  // namespace mcpi
