/*
 * Coulomb energy of an AO density from the oracle's ERIs:
 *   E_J = 1/2 sum_{pqrs} D_pq (pq|rs) D_rs.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no such
 * routine; this is the checker of the library's differentiation form, whose
 * nuclear gradient the tests compare with central differences of this energy.
 * It sums over shell quartets (ab|cd) with sa >= sb, sc >= sd and pair(ab) >=
 * pair(cd), weighting each by its number of permutational images, with the
 * ERIs of or_eri_quartet (eri.c, the Auto2e restatement).
 *
 * Pinned by tests/test_oracle_coulomb.py: equality with a plain numpy
 * contraction over every ordered shell quartet (no symmetry used), and the
 * closed form of one normalized s function, E_J = sqrt(alpha/pi) for D = 1.
 */
#include <stdlib.h>
#include "oracle.h"

double or_coulomb_energy(const or_basis *b, const double *D, int64_t ldd, double kappa)
{
    const int64_t nsh = b->nshells;
    int64_t *loc = malloc(sizeof(int64_t) * (size_t)nsh);
    int64_t nmax = 1;
    for (int64_t s = 0, l = 0; s < nsh; ++s) {
        loc[s] = l;
        const int64_t n = or_nang(b, s);
        l += n;
        if (n > nmax) nmax = n;
    }
    const int64_t np = nsh * (nsh + 1) / 2;
    int64_t *pa = malloc(sizeof(int64_t) * (size_t)np), *pb = malloc(sizeof(int64_t) * (size_t)np);
    for (int64_t sa = 0, k = 0; sa < nsh; ++sa)
        for (int64_t sb = 0; sb <= sa; ++sb, ++k) {
            pa[k] = sa;
            pb[k] = sb;
        }
    double E = 0.0;
#pragma omp parallel reduction(+ : E)
    {
        double *V = malloc(sizeof(double) * (size_t)(nmax * nmax * nmax * nmax));
#pragma omp for schedule(dynamic, 4)
        for (int64_t ij = 0; ij < np; ++ij) {
            const int64_t sa = pa[ij], sb = pb[ij];
            const int64_t na = or_nang(b, sa), nb = or_nang(b, sb);
            for (int64_t kl = 0; kl <= ij; ++kl) {
                const int64_t sc = pa[kl], sd = pb[kl];
                const int64_t nc = or_nang(b, sc), nd = or_nang(b, sd);
                or_eri_quartet(b, sa, sb, sc, sd, kappa, V);
                double s = 0.0;
                for (int64_t d = 0; d < nd; ++d)
                    for (int64_t c = 0; c < nc; ++c) {
                        const double dcd = D[(loc[sc] + c) + ldd * (loc[sd] + d)];
                        for (int64_t bb = 0; bb < nb; ++bb)
                            for (int64_t a = 0; a < na; ++a)
                                s += D[(loc[sa] + a) + ldd * (loc[sb] + bb)] * dcd *
                                     V[a + na * (bb + nb * (c + nc * d))];
                    }
                const double deg = (sa != sb ? 2.0 : 1.0) * (sc != sd ? 2.0 : 1.0) * (ij != kl ? 2.0 : 1.0);
                E += deg * s;
            }
        }
        free(V);
    }
    free(loc);
    free(pa);
    free(pb);
    return 0.5 * E;
}
