// Host test of pair_factorize (mcpi_amd/csrc/pair_factor.h): the
// Lieb-Liniger pair matrix factorizes into one rank-one block per total
// momentum; random rank-one blocks over a random partition are found and
// reproduced; a perturbed entry, an off-diagonal one-body entry, a rank-two
// block and an all-zero matrix are rejected; pairs without a term get no
// block.  Prints one line per case and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "pair_factor.h"

struct Z {
    double re, im;
};

static int fails = 0;
#define CHECK(c, msg)                                          \
    do {                                                       \
        if (!(c)) {                                            \
            std::printf("FAIL %s (line %d)\n", msg, __LINE__); \
            fails++;                                           \
        }                                                      \
    } while (0)

static int pair_index(int c1, int c2) { return c2 * (c2 + 1) / 2 + c1; }   // c1 <= c2

struct Mats {
    int M, NP, PWS, HS;
    std::vector<Z> pw, ph1;
    Mats(int m) : M(m), NP(m * (m + 1) / 2), PWS((m * (m + 1) / 2 + 15) / 16 * 16), HS(m <= 16 ? 16 : 32)
    {
        pw.assign((size_t)PWS * PWS, Z{0, 0});
        ph1.assign((size_t)HS * HS, Z{0, 0});
    }
    Z &W(int P, int Q) { return pw[(size_t)Q * PWS + P]; }
};

// max |W - za zb^T| over all entries, with zero outside the blocks
static double residual(Mats &m, const mcpi::PairFactor<Z> &f)
{
    double r = 0;
    for (int P = 0; P < m.NP; P++)
        for (int Q = 0; Q < m.NP; Q++) {
            double fr = 0, fi = 0;
            if (f.zblk[P] >= 0 && f.zblk[P] == f.zblk[Q]) {
                fr = f.za[P].re * f.zb[Q].re - f.za[P].im * f.zb[Q].im;
                fi = f.za[P].re * f.zb[Q].im + f.za[P].im * f.zb[Q].re;
            }
            r = std::fmax(r, std::hypot(m.W(P, Q).re - fr, m.W(P, Q).im - fi));
        }
    return r;
}

static void check_lists(const Mats &m, const mcpi::PairFactor<Z> &f)
{
    const int R = (int)f.zoff.size() - 1;
    CHECK(f.zoff[0] == 0, "zoff starts at 0");
    int listed = 0;
    for (int r = 0; r < R; r++)
        for (int j = f.zoff[r]; j < f.zoff[r + 1]; j++) {
            CHECK(f.zblk[f.zord[j]] == r, "zord entry in its block");
            listed++;
        }
    int with = 0;
    for (int P = 0; P < m.NP; P++) with += f.zblk[P] >= 0;
    CHECK(listed == with, "every pair with a block listed once");
}

int main()
{
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    for (int M : {3, 5, 12, 21}) {
        // Lieb-Liniger: W = g/2 m_P m_Q on equal total momentum, h1 diagonal
        Mats m(M);
        const double g = 3.5;
        for (int a2 = 0; a2 < M; a2++)
            for (int a1 = 0; a1 <= a2; a1++)
                for (int b2 = 0; b2 < M; b2++)
                    for (int b1 = 0; b1 <= b2; b1++)
                        if (a1 + a2 == b1 + b2)
                            m.W(pair_index(a1, a2), pair_index(b1, b2)) =
                                Z{g / 2 * (a1 == a2 ? 1 : 2) * (b1 == b2 ? 1 : 2), 0};
        for (int k = 0; k < M; k++) m.ph1[(size_t)k * m.HS + k] = Z{0.5 * k * k, 0};
        mcpi::PairFactor<Z> f;
        const bool ok = mcpi::pair_factorize(M, m.NP, m.PWS, m.HS, m.pw.data(), m.ph1.data(), f);
        CHECK(ok, "Lieb-Liniger factorizes");
        if (ok) {
            CHECK((int)f.zoff.size() - 1 == 2 * M - 1, "one block per total momentum");
            CHECK(residual(m, f) <= 1e-14 * g, "Lieb-Liniger reproduced");
            check_lists(m, f);
        }
        std::printf("M=%d Lieb-Liniger: %s, %d blocks\n", M, ok ? "factorized" : "rejected",
                    ok ? (int)f.zoff.size() - 1 : 0);
        // an off-diagonal one-body entry: rejected
        Mats m2 = m;
        m2.ph1[(size_t)1 * m2.HS + 0] = Z{1e-3, 0};
        CHECK(!mcpi::pair_factorize(M, m2.NP, m2.PWS, m2.HS, m2.pw.data(), m2.ph1.data(), f), "off-diagonal h1");
        // one entry perturbed by 1e-9 of the scale: rejected
        Mats m3 = m;
        m3.W(pair_index(0, M - 1), pair_index(0, M - 1)).re += 1e-9 * g;
        CHECK(!mcpi::pair_factorize(M, m3.NP, m3.PWS, m3.HS, m3.pw.data(), m3.ph1.data(), f), "perturbed entry");
    }
    for (int M : {4, 9, 12}) {
        // random partition of the pairs into blocks of random complex rank-one
        // matrices, some pairs without any term
        Mats m(M);
        std::vector<int> part(m.NP);
        for (int P = 0; P < m.NP; P++) part[P] = (int)(rng() % 7) - 1;   // -1: no term
        std::vector<Z> a(m.NP), b(m.NP);
        for (int P = 0; P < m.NP; P++) a[P] = Z{nd(rng), nd(rng)}, b[P] = Z{nd(rng), nd(rng)};
        for (int P = 0; P < m.NP; P++)
            for (int Q = 0; Q < m.NP; Q++)
                if (part[P] >= 0 && part[P] == part[Q])
                    m.W(P, Q) = Z{a[P].re * b[Q].re - a[P].im * b[Q].im, a[P].re * b[Q].im + a[P].im * b[Q].re};
        mcpi::PairFactor<Z> f;
        const bool ok = mcpi::pair_factorize(M, m.NP, m.PWS, m.HS, m.pw.data(), m.ph1.data(), f);
        CHECK(ok, "random rank-one blocks factorize");
        if (ok) {
            CHECK(residual(m, f) <= 1e-13 * 10, "random blocks reproduced");
            for (int P = 0; P < m.NP; P++) CHECK((part[P] < 0) == (f.zblk[P] < 0), "pairs without a term");
            check_lists(m, f);
        }
        std::printf("M=%d random rank-one blocks: %s, %d blocks\n", M, ok ? "factorized" : "rejected",
                    ok ? (int)f.zoff.size() - 1 : 0);
        // a rank-two block: rejected
        int p0 = -1, p1 = -1;
        for (int P = 0; P < m.NP && p1 < 0; P++)
            if (part[P] >= 0) {
                if (p0 < 0) p0 = P;
                else if (part[P] == part[p0]) p1 = P;
            }
        if (p1 >= 0) {
            Mats m2 = m;
            m2.W(p0, p0).re += 1.0;
            CHECK(!mcpi::pair_factorize(M, m2.NP, m2.PWS, m2.HS, m2.pw.data(), m2.ph1.data(), f), "rank two");
        }
    }
    {
        Mats m(6);   // no term at all
        mcpi::PairFactor<Z> f;
        CHECK(!mcpi::pair_factorize(6, m.NP, m.PWS, m.HS, m.pw.data(), m.ph1.data(), f), "all-zero matrix");
    }
    std::printf(fails ? "FAILED %d\n" : "all passed\n", fails);
    return fails ? 1 : 0;
}
