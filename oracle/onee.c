/*
 * One-electron integrals: the bare-nucleus Hamiltonian H0 = T + V_ne of
 * ints1e_gammcor_H0 (OneElectronInts_Gammcor.f90:13-73) and the overlap of
 * ints1e_S (OneElectronInts.f90:915-932).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates OneElectronInts.f90:
 * for every shell pair ShA >= ShB (ints1e_decode_pq, :194-239) and primitive
 * pair, the two-centre Hermite coefficients E^{ij}_t (ints1e_eijmatrix,
 * Helgaker 9.5.6-7, :38-178) with unit seed and exp(-mu AB^2) in the
 * prefactor; the overlap (pi/p)^{3/2} E^x_0 E^y_0 E^z_0 (:242-329); the
 * kinetic energy by Ferman & Valeev eq. 4.8 with the ket exponent
 * (ints1e_Kinetic_core, :715-805); the nuclear attraction
 * -sum_C Z_C (2 pi/p) sum_tuv E_t E_u E_v R_tuv(p, P - C), Helgaker 9.9.32
 * (ints1e_Coulomb, :518-712, point charges included).  Blocks of
 * unnormalized Cartesian primitives are moved to the AO basis with U and the
 * AO normalization factors (SpherGTO_TransformMatrix_U, sphergto.f90:47-81,
 * where NormFactorsCart cancels) and mirrored to the upper triangle
 * (real_smfill).
 *
 * Pinned by tests/test_oracle_onee.py: S and T against Gauss-Hermite
 * quadrature (exact for polynomial-times-Gaussian integrands) for s..f
 * shells on two centres, unit diagonal of S, closed forms for s functions
 * (overlap, kinetic 3a/2 at one centre, nuclear attraction F_0), and V for
 * p, d, f against the oracle's ERIs with a very tight s function at the
 * charge, whose square tends to a unit point charge.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define OR_PI 3.141592653589793

void or_rtuv(double *R, double *tmp, int64_t L, const double *F, double alpha, double X, double Y, double Z);

static int nc_(int l) { return (l + 1) * (l + 2) / 2; }

static void comp_(int l, int i, int *x, int *y, int *z)
{
    int k = 0;
    for (int a = l; a >= 0; --a)
        for (int b = l - a; b >= 0; --b) {
            if (k == i) { *x = a; *y = b; *z = l - a - b; return; }
            ++k;
        }
}

/* E(i, j, t), i <= imax, j <= jmax, t <= imax + jmax, unit seed */
static void e2c(double *E, int imax, int jmax, double p, double xpa, double xpb)
{
    const int tn = imax + jmax + 1, sj = tn, si = (jmax + 1) * tn;
    memset(E, 0, sizeof(double) * (size_t)((imax + 1) * si));
    const double i2p = 0.5 / p;
    E[0] = 1.0;
    for (int i = 0; i < imax; ++i)
        for (int t = 0; t <= i + 1; ++t) {
            double v = (t <= i ? xpa * E[i * si + t] : 0.0);
            if (t >= 1) v += i2p * E[i * si + t - 1];
            if (t + 1 <= i) v += (t + 1) * E[i * si + t + 1];
            E[(i + 1) * si + t] = v;
        }
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j < jmax; ++j)
            for (int t = 0; t <= i + j + 1; ++t) {
                const int o = i * si + j * sj;
                double v = (t <= i + j ? xpb * E[o + t] : 0.0);
                if (t >= 1) v += i2p * E[o + t - 1];
                if (t + 1 <= i + j) v += (t + 1) * E[o + t + 1];
                E[o + sj + t] = v;
            }
}

void or_one_electron(const or_basis *b, const double *atom_charges, int64_t npc, const double *pc_xyz,
                     const double *pc_q, double *S, double *T, double *V, int64_t ldh)
{
    const int64_t nsh = b->nshells;
    int64_t *loc = malloc(sizeof(int64_t) * (size_t)nsh);
    for (int64_t s = 0, l = 0; s < nsh; ++s) {
        loc[s] = l;
        l += or_nang(b, s);
    }
    /* charges: atoms, then point charges */
    const int64_t nq = b->natoms + npc;
    double *Q = malloc(sizeof(double) * 4 * (size_t)(nq > 0 ? nq : 1));
    for (int64_t a = 0; a < b->natoms; ++a) {
        Q[4 * a] = b->coords[3 * a];
        Q[4 * a + 1] = b->coords[3 * a + 1];
        Q[4 * a + 2] = b->coords[3 * a + 2];
        Q[4 * a + 3] = atom_charges ? atom_charges[a] : 0.0;
    }
    for (int64_t c = 0; c < npc; ++c) {
        double *q = Q + 4 * (b->natoms + c);
        q[0] = pc_xyz[3 * c];
        q[1] = pc_xyz[3 * c + 1];
        q[2] = pc_xyz[3 * c + 2];
        q[3] = pc_q[c];
    }
    const int64_t np = nsh * (nsh + 1) / 2;
#pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < np; ++k) {
        int64_t sa = 0, sb = 0;
        {   /* k -> (sa >= sb) */
            int64_t acc = 0;
            while (acc + sa + 1 <= k) { acc += sa + 1; ++sa; }
            sb = k - acc;
        }
        const int64_t pa = b->shell_params[sa] - 1, pb = b->shell_params[sb] - 1;
        const int la = (int)b->momentum[pa], lb = (int)b->momentum[pb];
        const double *A = b->coords + 3 * (b->shell_centers[sa] - 1);
        const double *B = b->coords + 3 * (b->shell_centers[sb] - 1);
        const int nca = nc_(la), ncb = nc_(lb), L = la + lb;
        double XS[100], XT[100], XV[100];
        memset(XS, 0, sizeof XS);
        memset(XT, 0, sizeof XT);
        memset(XV, 0, sizeof XV);
        const int imax = la, jmax = lb + 2, tn = imax + jmax + 1, sj = tn, si = (jmax + 1) * tn;
        double Ex[4 * 6 * 9], Ey[4 * 6 * 9], Ez[4 * 6 * 9];
        const int n1 = L + 1;
        double *R = malloc(sizeof(double) * (size_t)(n1 * n1 * n1)), *tmp = malloc(sizeof(double) * (size_t)(n1 * n1 * n1));
        double *RR = malloc(sizeof(double) * (size_t)(n1 * n1 * n1));
        double F[32];
        const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
        for (int jb = 0; jb < b->nprim[pb]; ++jb)
            for (int ia = 0; ia < b->nprim[pa]; ++ia) {
                const double a = b->exps[b->max_nprim * pa + ia], bb = b->exps[b->max_nprim * pb + jb];
                const double p = a + bb, mu = a * bb / p;
                const double P[3] = {(a * A[0] + bb * B[0]) / p, (a * A[1] + bb * B[1]) / p, (a * A[2] + bb * B[2]) / p};
                const double pref = b->cntr[b->max_nprim * pa + ia] * b->cntr[b->max_nprim * pb + jb] * exp(-mu * AB2);
                e2c(Ex, imax, jmax, p, P[0] - A[0], P[0] - B[0]);
                e2c(Ey, imax, jmax, p, P[1] - A[1], P[1] - B[1]);
                e2c(Ez, imax, jmax, p, P[2] - A[2], P[2] - B[2]);
                memset(RR, 0, sizeof(double) * (size_t)(n1 * n1 * n1));
                if (V)
                    for (int64_t c = 0; c < nq; ++c) {
                        const double *q = Q + 4 * c;
                        if (q[3] == 0.0) continue;
                        const double X = P[0] - q[0], Y = P[1] - q[1], Z = P[2] - q[2];
                        or_boys_fm(L, p * (X * X + Y * Y + Z * Z), F);
                        or_rtuv(R, tmp, L, F, p, X, Y, Z);
                        for (int e = 0; e < n1 * n1 * n1; ++e) RR[e] -= q[3] * R[e];
                    }
                const double sp = pref * pow(OR_PI / p, 1.5), vp = pref * 2.0 * OR_PI / p;
                for (int cb = 0; cb < ncb; ++cb)
                    for (int ca = 0; ca < nca; ++ca) {
                        int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
                        comp_(la, ca, &ax, &ay, &az);
                        comp_(lb, cb, &bx, &by, &bz);
#define EE(E_, i, j, t) (E_)[(i) * si + (j) * sj + (t)]
                        const double x0 = EE(Ex, ax, bx, 0), y0 = EE(Ey, ay, by, 0), z0 = EE(Ez, az, bz, 0);
                        const double sab = x0 * y0 * z0;
                        double Ix = (2 * bx + 1) * bb * sab - 2 * bb * bb * EE(Ex, ax, bx + 2, 0) * y0 * z0;
                        if (bx >= 2) Ix -= bx * (bx - 1) / 2.0 * EE(Ex, ax, bx - 2, 0) * y0 * z0;
                        double Iy = (2 * by + 1) * bb * sab - 2 * bb * bb * EE(Ey, ay, by + 2, 0) * x0 * z0;
                        if (by >= 2) Iy -= by * (by - 1) / 2.0 * EE(Ey, ay, by - 2, 0) * x0 * z0;
                        double Iz = (2 * bz + 1) * bb * sab - 2 * bb * bb * EE(Ez, az, bz + 2, 0) * x0 * y0;
                        if (bz >= 2) Iz -= bz * (bz - 1) / 2.0 * EE(Ez, az, bz - 2, 0) * x0 * y0;
                        double v = 0.0;
                        for (int t = 0; t <= ax + bx; ++t)
                            for (int u = 0; u <= ay + by; ++u)
                                for (int w = 0; w <= az + bz; ++w)
                                    v += EE(Ex, ax, bx, t) * EE(Ey, ay, by, u) * EE(Ez, az, bz, w) *
                                         RR[t + n1 * (u + n1 * w)];
#undef EE
                        XS[ca + nca * cb] += sp * sab;
                        XT[ca + nca * cb] += sp * (Ix + Iy + Iz);
                        XV[ca + nca * cb] += vp * v;
                    }
            }
        free(R);
        free(tmp);
        free(RR);
        /* the Cartesian diagonal block is symmetrized from its lower triangle */
        if (sa == sb)
            for (int cb = 0; cb < ncb; ++cb)
                for (int ca = 0; ca < cb; ++ca) {
                    XS[ca + nca * cb] = XS[cb + nca * ca];
                    XT[ca + nca * cb] = XT[cb + nca * ca];
                    XV[ca + nca * cb] = XV[cb + nca * ca];
                }
        /* AO basis: n(m) sum_c U(m, c) X (U = 1 for Cartesian AOs and l < 2) */
        const int ta = b->spher && la >= 2, tb = b->spher && lb >= 2;
        const int na = ta ? 2 * la + 1 : nca, nb = tb ? 2 * lb + 1 : ncb;
        double Ua[7 * 10], Ub[7 * 10];
        if (ta) or_spher_u(la, Ua);
        if (tb) or_spher_u(lb, Ub);
        const double *nA = b->norm + b->max_nang * pa, *nB = b->norm + b->max_nang * pb;
        double *outs[3] = {S, T, V};
        const double *X[3] = {XS, XT, XV};
        for (int o = 0; o < 3; ++o) {
            if (!outs[o]) continue;
            for (int mb = 0; mb < nb; ++mb)
                for (int ma = 0; ma < na; ++ma) {
                    if (sa == sb && ma < mb) continue;
                    double acc = 0.0;
                    for (int cb = 0; cb < ncb; ++cb) {
                        const double ub = tb ? Ub[mb * ncb + cb] : (mb == cb ? 1.0 : 0.0);
                        if (ub == 0.0) continue;
                        for (int ca = 0; ca < nca; ++ca) {
                            const double ua = ta ? Ua[ma * nca + ca] : (ma == ca ? 1.0 : 0.0);
                            if (ua != 0.0) acc += ua * ub * X[o][ca + nca * cb];
                        }
                    }
                    const double val = nA[ma] * nB[mb] * acc;
                    outs[o][(loc[sa] + ma) + ldh * (loc[sb] + mb)] = val;
                    outs[o][(loc[sb] + mb) + ldh * (loc[sa] + ma)] = val;
                }
        }
    }
    free(loc);
    free(Q);
}
