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

/* Nuclear gradient of E_J from derivative ERIs (or_eri_deriv_a, eri.c):
 *   dE_J/dR_A = 2 sum_{a on A} sum_b sum_{cd} D_ab D_cd (d_A a b|cd),
 * E_J being symmetric in the two pairs and D symmetric, so only the first
 * function of the bra is differentiated.  Every ordered bra shell pair whose
 * first shell sits on an atom of `atoms` (all atoms when natoms_sel < 0),
 * ket shell pairs sc >= sd weighted 2 when sc != sd.  G is (3, natoms),
 * zeroed here; atoms outside the selection stay zero.  Returns non-zero if
 * a derivative class is beyond the oracle's tables. */
int or_coulomb_grad(const or_basis *b, const double *D, int64_t ldd, double kappa, int64_t natoms_sel,
                    const int64_t *atoms, double *G)
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
    char *want = calloc((size_t)b->natoms, 1);
    for (int64_t i = 0; i < b->natoms; ++i) want[i] = natoms_sel < 0;
    for (int64_t i = 0; i < natoms_sel; ++i)
        if (atoms[i] >= 0 && atoms[i] < b->natoms) want[atoms[i]] = 1;
    for (int64_t i = 0; i < 3 * b->natoms; ++i) G[i] = 0.0;
    int64_t nbra = 0;
    int64_t *ba = malloc(sizeof(int64_t) * (size_t)(nsh * nsh)), *bb = malloc(sizeof(int64_t) * (size_t)(nsh * nsh));
    for (int64_t sa = 0; sa < nsh; ++sa)
        if (want[b->shell_centers[sa] - 1])
            for (int64_t sb = 0; sb < nsh; ++sb) {
                ba[nbra] = sa;
                bb[nbra] = sb;
                ++nbra;
            }
    int status = 0;
#pragma omp parallel
    {
        double *dV = malloc(sizeof(double) * (size_t)(3 * nmax * nmax * nmax * nmax));
        double *Gl = calloc((size_t)(3 * b->natoms), sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < nbra; ++k) {
            const int64_t sa = ba[k], sb = bb[k];
            const int64_t na = or_nang(b, sa), nbn = or_nang(b, sb);
            double g[3] = {0.0, 0.0, 0.0};
            for (int64_t sc = 0; sc < nsh; ++sc)
                for (int64_t sd = 0; sd <= sc; ++sd) {
                    const int64_t nc = or_nang(b, sc), nd = or_nang(b, sd);
                    if (or_eri_deriv_a(b, sa, sb, sc, sd, kappa, dV) != 0) {
#pragma omp atomic write
                        status = 1;
                        continue;
                    }
                    const int64_t nq = na * nbn * nc * nd;
                    const double w = sc != sd ? 2.0 : 1.0;
                    for (int dir = 0; dir < 3; ++dir) {
                        double s = 0.0;
                        for (int64_t d = 0; d < nd; ++d)
                            for (int64_t c = 0; c < nc; ++c) {
                                const double dcd = D[(loc[sc] + c) + ldd * (loc[sd] + d)];
                                for (int64_t q = 0; q < nbn; ++q)
                                    for (int64_t a = 0; a < na; ++a)
                                        s += D[(loc[sa] + a) + ldd * (loc[sb] + q)] * dcd *
                                             dV[dir * nq + a + na * (q + nbn * (c + nc * d))];
                            }
                        g[dir] += w * s;
                    }
                }
            const int64_t at = b->shell_centers[sa] - 1;
            for (int dir = 0; dir < 3; ++dir) Gl[3 * at + dir] += 2.0 * g[dir];
        }
#pragma omp critical
        for (int64_t i = 0; i < 3 * b->natoms; ++i) G[i] += Gl[i];
        free(Gl);
        free(dV);
    }
    free(ba);
    free(bb);
    free(want);
    free(loc);
    return status;
}
