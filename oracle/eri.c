/*
 * Oracle: electron-repulsion integrals (ab|cd) over contracted Gaussian
 * shells, restating the reference's generated Auto2e engine.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Algorithm (as the reference):
 *   canonical shell order A>=B, C>=D, AB>=CD          Auto2e.py:1612-1647
 *   W(e0|f0), e=la..la+lb, f=lc..lc+ld, one block per (e,f), each block with
 *   its own Boys evaluation at top order e+f           auto2e_WMatrix_part1.f90:99-148,445-506
 *     primitive loop order Gd,Gc,Gb,Ga; prefactors P,Q,mu,K,alpha,N,E000
 *                                                      auto2e_WMatrix_part1.f90:118-140
 *     one-centre Hermite coefficients E^{i,0}_t        auto2e_Hermite.f90:10-60
 *     R_tuv from (-2 alpha)^n F_n (Helgaker 9.9.18-20)  Auto2e.py:80-119
 *     ket transform with (-1)^(tau+nu+phi), bra transform
 *                                                      auto2e_WMatrix_part1.f90:280-298
 *   HGP horizontal recurrence on CD then AB            Auto2e.py:137-177, 298-338
 *   Cartesian -> real solid harmonics for l>=2         AutoSpherTransf.py:62-158,
 *                                                      Auto2eGlobal.py:21-26
 *   normalization NormA*NormB*NormC*NormD              Auto2e.py:1287-1367
 *
 * Pinned by tests/test_oracle_eri.py: closed-form (ss|ss) (also with the
 * erf-attenuated Kappa>0 operator), an independent Obara-Saika/HGP
 * implementation in mpmath with exact Boys values (different recurrences,
 * different centres), 8-fold permutational symmetry, translation invariance,
 * rotation invariance of shell-block Frobenius norms, unit self-overlap
 * of the normalized functions, and Schwarz bounds.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define OR_PI 3.141592653589793

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* position of (lx,ly,lz) in the order lx=l..0, ly=l-lx..0
 * (Auto2eGlobal.py:35-43, basis_sets.f90:797-807) */
static int cart_idx(int lx, int ly, int lz)
{
    int l = lx + ly + lz;
    return (l - lx) * (l - lx + 1) / 2 + (l - lx - ly);
}

static void cart_comp(int l, int i, int *lx, int *ly, int *lz)
{
    int k = 0;
    *lx = *ly = *lz = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (k == i) { *lx = x; *ly = y; *lz = l - x - y; return; }
            ++k;
        }
}

/* ---------------- real solid harmonics (AutoSpherTransf.py:62-158) ------- */
static double fact(int n) { double r = 1.0; for (int i = 2; i <= n; ++i) r *= i; return r; }
static double binom(int n, int k)
{
    if (k < 0 || k > n || n < 0) return 0.0;
    return fact(n) / (fact(k) * fact(n - k));
}

void or_spher_u(int64_t l, double *U)
{
    int L = (int)l, nc = ncart(L);
    memset(U, 0, sizeof(double) * (2 * L + 1) * nc);
    for (int m = -L; m <= L; ++m) {
        int am = abs(m), vm = (m >= 0) ? 0 : 1;
        /* Nslm: unit-sphere normalized, AutoSpherTransf.py:62-85 */
        double npref = sqrt((2.0 * L + 1.0) / (4.0 * OR_PI)) * sqrt(fact(L + am)) *
                       sqrt(fact(L - am)) * (m == 0 ? 1.0 : sqrt(2.0)) /
                       (pow(2.0, am) * fact(L));
        for (int t = 0; t <= (L - am) / 2; ++t)
            for (int u = 0; u <= t; ++u)
                for (int k = 0; k <= (am - vm) / 2; ++k) {
                    int v = 2 * k + vm;
                    int lx = 2 * t + am - 2 * u - v, ly = 2 * u + v, lz = L - 2 * t - am;
                    /* Clmtuv, AutoSpherTransf.py:88-110 */
                    double a = binom(L, t) * binom(L - t, am + t);
                    double b = binom(t, u) * binom(am, v);
                    int sgnexp = t + (v - vm) / 2;
                    double c = ((sgnexp % 2) ? -1.0 : 1.0) * pow(4.0, t);
                    U[(m + L) * nc + cart_idx(lx, ly, lz)] += npref * (a * b / c);
                }
    }
}

/* ---------------- shells ---------------- */
typedef struct {
    const double *R;
    int l, np;
    const double *c, *e, *nrm;
} shell_t;

static void get_shell(const or_basis *b, int64_t s, shell_t *sh)
{
    int64_t p = b->shell_params[s] - 1, a = b->shell_centers[s] - 1;
    sh->R = b->coords + 3 * a;
    sh->l = (int)b->momentum[p];
    sh->np = (int)b->nprim[p];
    sh->c = b->cntr + b->max_nprim * p;
    sh->e = b->exps + b->max_nprim * p;
    sh->nrm = b->norm + b->max_nang * p;
}

int64_t or_nang(const or_basis *b, int64_t s)
{
    int64_t l = b->momentum[b->shell_params[s] - 1];
    return b->spher ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
}

/* one-centre Hermite coefficients E[i][t], i<=imax (auto2e_Hermite.f90:10-60):
 * E^{i}_0 = X E^{i-1}_0 + E^{i-1}_1,  E^{i}_t = i/(2pt) E^{i-1}_{t-1} */
#define EMAX 11
static void etcoeffs(double E[EMAX][EMAX], int imax, double p, double X, double e000)
{
    E[0][0] = e000;
    for (int i = 1; i <= imax; ++i) {
        E[i][0] = X * E[i - 1][0] + (i - 1 >= 1 ? E[i - 1][1] : 0.0);
        for (int t = 1; t <= i; ++t) E[i][t] = (double)i / (2.0 * t * p) * E[i - 1][t - 1];
    }
}

/* R_tuv, t+u+v<=L (Helgaker 9.9.18-20; recurrence order t, then u, then v as
 * Auto2e.py:80-119).  R is (L+1)^3 dense, index t + (L+1)(u + (L+1) v). */
static void rtuv(double *R, double *tmp, int L, const double *F, double alpha,
                 double X, double Y, double Z)
{
    int n1 = L + 1;
#define IX(t, u, v) ((t) + n1 * ((u) + n1 * (v)))
    double *A = R, *B = tmp;
    /* level n = L: only R_000 */
    double m2a = -2.0 * alpha, pw = 1.0;
    for (int n = 0; n < L; ++n) pw *= m2a;
    /* iterate from n=L down to 0; at level n the max t+u+v is L-n.  Use
     * ping-pong so that the final level (n=0) lands in R. */
    double *cur = ((L % 2) == 0) ? A : B, *prev = ((L % 2) == 0) ? B : A;
    cur[IX(0, 0, 0)] = pw * F[L];
    for (int n = L - 1; n >= 0; --n) {
        double *t_ = prev; prev = cur; cur = t_;
        pw = 1.0;
        for (int i = 0; i < n; ++i) pw *= m2a;   /* (-2 alpha)^n */
        int J = L - n;
        for (int v = 0; v <= J; ++v)
            for (int u = 0; u <= J - v; ++u)
                for (int t = 0; t <= J - v - u; ++t) {
                    double r;
                    if (t >= 1)
                        r = (t >= 2 ? (t - 1) * prev[IX(t - 2, u, v)] : 0.0) + X * prev[IX(t - 1, u, v)];
                    else if (u >= 1)
                        r = (u >= 2 ? (u - 1) * prev[IX(t, u - 2, v)] : 0.0) + Y * prev[IX(t, u - 1, v)];
                    else if (v >= 1)
                        r = (v >= 2 ? (v - 1) * prev[IX(t, u, v - 2)] : 0.0) + Z * prev[IX(t, u, v - 1)];
                    else
                        r = pw * F[n];
                    cur[IX(t, u, v)] = r;
                }
    }
#undef IX
}

/* exported for onee.c: R_tuv as rtuv(), R and tmp (L+1)^3 each */
void or_rtuv(double *R, double *tmp, int64_t L, const double *F, double alpha, double X, double Y, double Z)
{
    rtuv(R, tmp, (int)L, F, alpha, X, Y, Z);
}

/* horizontal recurrence (Head-Gordon & Pople eq. 18):
 * (c, d+1_i) = (c+1_i, d) + (C_i - D_i)(c, d), direction = first nonzero of d
 * in x,y,z order (Auto2e.py:137-177).
 * in: values for c-components with l in [l1, l1+l2], packed by l then cart_idx
 * out: out[c + ncart(l1) * d] */
static void hrr(const double *in, int l1, int l2, const double CD[3], double *out)
{
    int nc1 = ncart(l1);
    if (l2 == 0) { memcpy(out, in, sizeof(double) * nc1); return; }
    /* offsets of packed c-levels */
    int off[24];
    off[0] = 0;
    for (int k = 0; k <= l2; ++k) off[k + 1] = off[k] + ncart(l1 + k);
    int ntot = off[l2 + 1];
    /* G_prev[c (packed, levels l1..l1+l2-dl+1)][d (ncart(dl-1))] */
    double *prev = (double *)malloc(sizeof(double) * ntot);
    memcpy(prev, in, sizeof(double) * ntot);
    int prev_nd = 1;
    for (int dl = 1; dl <= l2; ++dl) {
        int nd = ncart(dl), ncp = off[l2 - dl + 1];  /* c levels l1..l1+l2-dl */
        int ncp_prev = off[l2 - dl + 2];
        double *cur = (double *)malloc(sizeof(double) * ncp * nd);
        for (int id = 0; id < nd; ++id) {
            int dx, dy, dz;
            cart_comp(dl, id, &dx, &dy, &dz);
            int dir = dx > 0 ? 0 : (dy > 0 ? 1 : 2);
            int ex = dx - (dir == 0), ey = dy - (dir == 1), ez = dz - (dir == 2);
            int idm = cart_idx(ex, ey, ez);
            for (int lc = 0; lc <= l2 - dl; ++lc) {
                int l = l1 + lc;
                for (int ic = 0; ic < ncart(l); ++ic) {
                    int cx, cy, cz;
                    cart_comp(l, ic, &cx, &cy, &cz);
                    int px = cx + (dir == 0), py = cy + (dir == 1), pz = cz + (dir == 2);
                    int ip = off[lc + 1] + cart_idx(px, py, pz);
                    int i0 = off[lc] + ic;
                    cur[i0 + ncp * id] = prev[ip + ncp_prev * idm] + CD[dir] * prev[i0 + ncp_prev * idm];
                }
            }
        }
        free(prev);
        prev = cur;
        prev_nd = nd;
    }
    (void)prev_nd;
    for (int id = 0; id < ncart(l2); ++id)
        for (int ic = 0; ic < nc1; ++ic) out[ic + nc1 * id] = prev[ic + off[1] * id];
    free(prev);
}

/* Cartesian (a,b,c,d) integrals for canonical shells, unnormalized functions
 * x^lx y^ly z^lz sum_i c_i exp(-a_i r^2).  H[a + na(b + nb(c + nc d))]. */
static void eri_cart_canonical(const shell_t *A, const shell_t *B, const shell_t *C,
                               const shell_t *D, double kappa, int single_boys, double *H)
{
    const int la = A->l, lb = B->l, lc = C->l, ld = D->l;
    const int Ltot = la + lb + lc + ld;
    int roff[12], coff[12];
    roff[0] = 0;
    for (int k = 0; k <= lb; ++k) roff[k + 1] = roff[k] + ncart(la + k);
    coff[0] = 0;
    for (int k = 0; k <= ld; ++k) coff[k + 1] = coff[k] + ncart(lc + k);
    const int nr = roff[lb + 1], ncol = coff[ld + 1];
    double *W = (double *)calloc((size_t)nr * ncol, sizeof(double));
    const int n1 = Ltot + 1;
    double *R = (double *)malloc(sizeof(double) * n1 * n1 * n1);
    double *Rt = (double *)malloc(sizeof(double) * n1 * n1 * n1);
    double *S = (double *)malloc(sizeof(double) * n1 * n1 * n1);
    double F[32];
    double Exa[EMAX][EMAX], Eya[EMAX][EMAX], Eza[EMAX][EMAX];
    double Exc[EMAX][EMAX], Eyc[EMAX][EMAX], Ezc[EMAX][EMAX];
    const double *Ra = A->R, *Rb = B->R, *Rc = C->R, *Rd = D->R;
    double Rab2 = (Ra[0] - Rb[0]) * (Ra[0] - Rb[0]) + (Ra[1] - Rb[1]) * (Ra[1] - Rb[1]) +
                  (Ra[2] - Rb[2]) * (Ra[2] - Rb[2]);
    double Rcd2 = (Rc[0] - Rd[0]) * (Rc[0] - Rd[0]) + (Rc[1] - Rd[1]) * (Rc[1] - Rd[1]) +
                  (Rc[2] - Rd[2]) * (Rc[2] - Rd[2]);
    const double twopi52 = 2.0 * pow(OR_PI, 2.5);

    for (int kb = 0; kb <= lb; ++kb)
        for (int kd = 0; kd <= ld; ++kd) {
            const int e = la + kb, f = lc + kd, Lef = e + f;
            const int mtop = single_boys ? Ltot : Lef;
            const int nR = Lef + 1;
            for (int gd = 0; gd < D->np; ++gd)
                for (int gc = 0; gc < C->np; ++gc) {
                    double ec = C->e[gc], ed = D->e[gd];
                    double Q = ec + ed, MuCD = ec * ed / Q;
                    double Rq[3], Rqc[3];
                    for (int i = 0; i < 3; ++i) { Rq[i] = (ec * Rc[i] + ed * Rd[i]) / Q; Rqc[i] = Rq[i] - Rc[i]; }
                    etcoeffs(Exc, f, Q, Rqc[0], 1.0);
                    etcoeffs(Eyc, f, Q, Rqc[1], 1.0);
                    etcoeffs(Ezc, f, Q, Rqc[2], 1.0);
                    for (int gb = 0; gb < B->np; ++gb)
                        for (int ga = 0; ga < A->np; ++ga) {
                            double ea = A->e[ga], eb = B->e[gb];
                            double P = ea + eb, MuAB = ea * eb / P;
                            double Rp[3], Rpq[3], Rpa[3];
                            for (int i = 0; i < 3; ++i) {
                                Rp[i] = (ea * Ra[i] + eb * Rb[i]) / P;
                                Rpq[i] = Rp[i] - Rq[i];
                                Rpa[i] = Rp[i] - Ra[i];
                            }
                            double K = exp(-MuAB * Rab2 - MuCD * Rcd2);
                            double PQ = P * Q;
                            double alpha = PQ / (P + Q + kappa * PQ);
                            double N = twopi52 / (PQ * sqrt(P + Q + kappa * PQ));
                            double E000 = N * A->c[ga] * B->c[gb] * C->c[gc] * D->c[gd] * K;
                            etcoeffs(Exa, e, P, Rpa[0], E000);
                            etcoeffs(Eya, e, P, Rpa[1], 1.0);
                            etcoeffs(Eza, e, P, Rpa[2], 1.0);
                            double x = alpha * (Rpq[0] * Rpq[0] + Rpq[1] * Rpq[1] + Rpq[2] * Rpq[2]);
                            or_boys_fm(mtop, x, F);
                            rtuv(R, Rt, Lef, F, alpha, Rpq[0], Rpq[1], Rpq[2]);
#define RI(t, u, v) R[(t) + nR * ((u) + nR * (v))]
#define SI(t, u, v) S[(t) + (e + 1) * ((u) + (e + 1) * (v))]
                            for (int kc = 0; kc < ncart(f); ++kc) {
                                int kx, ky, kz;
                                cart_comp(f, kc, &kx, &ky, &kz);
                                /* ket transform */
                                for (int v = 0; v <= e; ++v)
                                    for (int u = 0; u <= e - v; ++u)
                                        for (int t = 0; t <= e - v - u; ++t) {
                                            double s = 0.0;
                                            for (int ph = 0; ph <= kz; ++ph)
                                                for (int nu = 0; nu <= ky; ++nu)
                                                    for (int ta = 0; ta <= kx; ++ta) {
                                                        double sg = ((ta + nu + ph) & 1) ? -1.0 : 1.0;
                                                        s += sg * Exc[kx][ta] * Eyc[ky][nu] * Ezc[kz][ph] *
                                                             RI(t + ta, u + nu, v + ph);
                                                    }
                                            SI(t, u, v) = s;
                                        }
                                /* bra transform */
                                for (int ic = 0; ic < ncart(e); ++ic) {
                                    int ex, ey, ez;
                                    cart_comp(e, ic, &ex, &ey, &ez);
                                    double w = 0.0;
                                    for (int v = 0; v <= ez; ++v)
                                        for (int u = 0; u <= ey; ++u)
                                            for (int t = 0; t <= ex; ++t)
                                                w += Exa[ex][t] * Eya[ey][u] * Eza[ez][v] * SI(t, u, v);
                                    W[(roff[kb] + ic) + (size_t)nr * (coff[kd] + kc)] += w;
                                }
                            }
#undef RI
#undef SI
                        }
                }
        }
    free(R); free(Rt); free(S);
    /* ket HRR: per W row, columns (f-comps) -> (c,d) */
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    double CD[3] = {Rc[0] - Rd[0], Rc[1] - Rd[1], Rc[2] - Rd[2]};
    double AB[3] = {Ra[0] - Rb[0], Ra[1] - Rb[1], Ra[2] - Rb[2]};
    double *T = (double *)malloc(sizeof(double) * nr * nc * nd);   /* T[row][cd] */
    double *vin = (double *)malloc(sizeof(double) * (nr > ncol ? nr : ncol));
    double *vout = (double *)malloc(sizeof(double) * (na * nb > nc * nd ? na * nb : nc * nd));
    for (int r = 0; r < nr; ++r) {
        for (int k = 0; k < ncol; ++k) vin[k] = W[r + (size_t)nr * k];
        hrr(vin, lc, ld, CD, vout);
        for (int k = 0; k < nc * nd; ++k) T[r + (size_t)nr * k] = vout[k];
    }
    /* bra HRR: per (c,d) column */
    for (int k = 0; k < nc * nd; ++k) {
        hrr(T + (size_t)nr * k, la, lb, AB, vout);
        for (int i = 0; i < na * nb; ++i) H[i + (size_t)na * nb * k] = vout[i];
    }
    free(vin); free(vout); free(T); free(W);
}

/* apply U (nout x nin) to index `which` of a 4-index array with dims n[4] */
static void transform_index(double *X, int n[4], int which, const double *U, int nout)
{
    int nin = n[which];
    int stride = 1;
    for (int k = 0; k < which; ++k) stride *= n[k];
    int outer = 1;
    for (int k = which + 1; k < 4; ++k) outer *= n[k];
    int total_out = stride * nout * outer;
    double *Y = (double *)calloc(total_out, sizeof(double));
    for (int o = 0; o < outer; ++o)
        for (int m = 0; m < nout; ++m)
            for (int c = 0; c < nin; ++c) {
                double u = U[m * nin + c];
                if (u == 0.0) continue;
                for (int s = 0; s < stride; ++s)
                    Y[s + stride * (m + nout * o)] += u * X[s + stride * (c + nin * o)];
            }
    memcpy(X, Y, sizeof(double) * total_out);
    free(Y);
    n[which] = nout;
}

static void eri_any(const or_basis *b, int64_t s1, int64_t s2, int64_t s3, int64_t s4,
                    double kappa, int single_boys, double *V)
{
    shell_t sh[4];
    int64_t ss[4] = {s1, s2, s3, s4};
    for (int k = 0; k < 4; ++k) get_shell(b, ss[k], &sh[k]);
    /* Canonical4Tuple, Auto2e.py:1612-1647 */
    int m[4] = {0, 1, 2, 3};
    int l[4] = {sh[0].l, sh[1].l, sh[2].l, sh[3].l};
    if (l[1] > l[0]) { int t = m[0]; m[0] = m[1]; m[1] = t; t = l[0]; l[0] = l[1]; l[1] = t; }
    if (l[3] > l[2]) { int t = m[2]; m[2] = m[3]; m[3] = t; t = l[2]; l[2] = l[3]; l[3] = t; }
    if (l[2] > l[0] || (l[2] == l[0] && l[3] > l[1])) {
        int t0 = m[0], t1 = m[1]; m[0] = m[2]; m[1] = m[3]; m[2] = t0; m[3] = t1;
        t0 = l[0]; t1 = l[1]; l[0] = l[2]; l[1] = l[3]; l[2] = t0; l[3] = t1;
    }
    int nn[4];
    for (int k = 0; k < 4; ++k) nn[k] = ncart(l[k]);
    int maxdim = 1;
    for (int k = 0; k < 4; ++k) maxdim *= nn[k];
    double *H = (double *)malloc(sizeof(double) * maxdim);
    eri_cart_canonical(&sh[m[0]], &sh[m[1]], &sh[m[2]], &sh[m[3]], kappa, single_boys, H);
    /* spherical transform for l >= 2 (SPHER_TRANSF_LMIN = 2) */
    if (b->spher) {
        for (int k = 0; k < 4; ++k)
            if (l[k] >= 2) {
                double U[11 * 66];
                or_spher_u(l[k], U);
                transform_index(H, nn, k, U, 2 * l[k] + 1);
            }
    }
    /* normalize */
    for (int i3 = 0; i3 < nn[3]; ++i3)
        for (int i2 = 0; i2 < nn[2]; ++i2)
            for (int i1 = 0; i1 < nn[1]; ++i1)
                for (int i0 = 0; i0 < nn[0]; ++i0)
                    H[i0 + nn[0] * (i1 + nn[1] * (i2 + nn[2] * i3))] *=
                        sh[m[0]].nrm[i0] * sh[m[1]].nrm[i1] * sh[m[2]].nrm[i2] * sh[m[3]].nrm[i3];
    /* scatter to the caller's order: position k of the canonical tuple is the
     * caller's shell m[k] */
    int no[4];
    for (int k = 0; k < 4; ++k) no[m[k]] = nn[k];
    int idx[4];
    for (idx[3] = 0; idx[3] < nn[3]; ++idx[3])
        for (idx[2] = 0; idx[2] < nn[2]; ++idx[2])
            for (idx[1] = 0; idx[1] < nn[1]; ++idx[1])
                for (idx[0] = 0; idx[0] < nn[0]; ++idx[0]) {
                    int o[4];
                    for (int k = 0; k < 4; ++k) o[m[k]] = idx[k];
                    V[o[0] + no[0] * (o[1] + no[1] * (o[2] + no[2] * o[3]))] =
                        H[idx[0] + nn[0] * (idx[1] + nn[1] * (idx[2] + nn[2] * idx[3]))];
                }
    free(H);
}

void or_eri_quartet(const or_basis *b, int64_t sa, int64_t sb, int64_t sc, int64_t sd,
                    double kappa, double *V)
{
    eri_any(b, sa, sb, sc, sd, kappa, 0, V);
}

void or_eri_quartet_single_boys(const or_basis *b, int64_t sa, int64_t sb, int64_t sc,
                                int64_t sd, double kappa, double *V)
{
    eri_any(b, sa, sb, sc, sd, kappa, 1, V);
}

/* Dense (pq|rs) over a list of shell pairs (0-based (A, B) with A >= B), the
 * orbital pairs of each in the reference's storage order (b outer, a inner,
 * a >= b when A == B; chol2_write_Vdiag, TwoStepCholesky_Step1.f90:175-199).
 * V is n x n column-major, n = sum of the pair dimensions.  Test helper for
 * dense restatements of the decomposition. */
void or_eri_pair_matrix(const or_basis *b, int64_t npairs, const int64_t *pairs, double kappa, double *V, int64_t n)
{
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (npairs + 1));
    off[0] = 0;
    for (int64_t i = 0; i < npairs; ++i) {
        int64_t na = or_nang(b, pairs[2 * i]), nb = or_nang(b, pairs[2 * i + 1]);
        off[i + 1] = off[i] + (pairs[2 * i] == pairs[2 * i + 1] ? na * (na + 1) / 2 : na * nb);
    }
    #pragma omp parallel for schedule(dynamic)
    for (int64_t j = 0; j < npairs; ++j) {
        const int64_t sc = pairs[2 * j], sd = pairs[2 * j + 1];
        const int64_t nc = or_nang(b, sc), nd = or_nang(b, sd);
        for (int64_t i = j; i < npairs; ++i) {
            const int64_t sa = pairs[2 * i], sb = pairs[2 * i + 1];
            const int64_t na = or_nang(b, sa), nb = or_nang(b, sb);
            double *blk = (double *)malloc(sizeof(double) * na * nb * nc * nd);
            or_eri_quartet(b, sa, sb, sc, sd, kappa, blk);
            int64_t q = off[j];
            for (int64_t d = 0; d < nd; ++d)
                for (int64_t c = (sc == sd ? d : 0); c < nc; ++c, ++q) {
                    int64_t p = off[i];
                    for (int64_t bb = 0; bb < nb; ++bb)
                        for (int64_t a = (sa == sb ? bb : 0); a < na; ++a, ++p) {
                            const double v = blk[a + na * (bb + nb * (c + nc * d))];
                            V[p + n * q] = v;
                            V[q + n * p] = v;
                        }
                }
            free(blk);
        }
    }
    free(off);
}

/* ---------------- derivative integrals ----------------
 * Not part of the reference (it has no gradients): the checker of the
 * library's nuclear gradient.  The derivative of a Cartesian Gaussian with
 * respect to its own centre is a combination of Gaussians one order up and
 * one order down on the same centre,
 *   d/dA_x [x_A^lx y_A^ly z_A^lz exp(-a r_A^2)]
 *       = 2a x_A^(lx+1) y_A^ly z_A^lz exp(-a r_A^2) - lx x_A^(lx-1) y_A^ly z_A^lz exp(-a r_A^2),
 * so (d_A a b|cd) comes from two ordinary Cartesian ERI blocks -- shell a
 * raised to la + 1 with coefficients 2 a_i c_i, and lowered to la - 1 with
 * coefficients c_i -- followed by the solid-harmonic transform and the
 * normalization of the original shells.  No Hermite-derivative identity is
 * used, so this is an independent route to the gradient. */

/* Cartesian block of four shells (unnormalized, no spherical transform) in
 * the caller's order H[i0 + n0 (i1 + n1 (i2 + n2 i3))], via the canonical
 * evaluation (Canonical4Tuple, Auto2e.py:1612-1647). */
static void eri_cart_callers(const shell_t *s0, const shell_t *s1, const shell_t *s2, const shell_t *s3,
                             double kappa, double *V)
{
    const shell_t *sh[4] = {s0, s1, s2, s3};
    int m[4] = {0, 1, 2, 3};
    int l[4] = {s0->l, s1->l, s2->l, s3->l};
    if (l[1] > l[0]) { int t = m[0]; m[0] = m[1]; m[1] = t; t = l[0]; l[0] = l[1]; l[1] = t; }
    if (l[3] > l[2]) { int t = m[2]; m[2] = m[3]; m[3] = t; t = l[2]; l[2] = l[3]; l[3] = t; }
    if (l[2] > l[0] || (l[2] == l[0] && l[3] > l[1])) {
        int t0 = m[0], t1 = m[1]; m[0] = m[2]; m[1] = m[3]; m[2] = t0; m[3] = t1;
        t0 = l[0]; t1 = l[1]; l[0] = l[2]; l[1] = l[3]; l[2] = t0; l[3] = t1;
    }
    int nn[4], no[4];
    for (int k = 0; k < 4; ++k) nn[k] = ncart(l[k]);
    for (int k = 0; k < 4; ++k) no[m[k]] = nn[k];
    double *H = (double *)malloc(sizeof(double) * nn[0] * nn[1] * nn[2] * nn[3]);
    eri_cart_canonical(sh[m[0]], sh[m[1]], sh[m[2]], sh[m[3]], kappa, 0, H);
    int idx[4];
    for (idx[3] = 0; idx[3] < nn[3]; ++idx[3])
        for (idx[2] = 0; idx[2] < nn[2]; ++idx[2])
            for (idx[1] = 0; idx[1] < nn[1]; ++idx[1])
                for (idx[0] = 0; idx[0] < nn[0]; ++idx[0]) {
                    int o[4];
                    for (int k = 0; k < 4; ++k) o[m[k]] = idx[k];
                    V[o[0] + no[0] * (o[1] + no[1] * (o[2] + no[2] * o[3]))] =
                        H[idx[0] + nn[0] * (idx[1] + nn[1] * (idx[2] + nn[2] * idx[3]))];
                }
    free(H);
}

int or_eri_deriv_a(const or_basis *b, int64_t sa, int64_t sb, int64_t sc, int64_t sd, double kappa, double *dV)
{
    shell_t sh[4];
    int64_t ss[4] = {sa, sb, sc, sd};
    for (int k = 0; k < 4; ++k) get_shell(b, ss[k], &sh[k]);
    const int la = sh[0].l;
    if (la + 1 + sh[1].l >= EMAX || sh[2].l + sh[3].l >= EMAX || la + 1 + sh[1].l + sh[2].l + sh[3].l > 20)
        return 1;
    double cp[64];
    if (sh[0].np > 64) return 2;
    for (int i = 0; i < sh[0].np; ++i) cp[i] = 2.0 * sh[0].e[i] * sh[0].c[i];
    shell_t up = sh[0], dn = sh[0];
    up.l = la + 1;
    up.c = cp;
    dn.l = la - 1;
    const int n0 = ncart(la), n0p = ncart(la + 1), n0m = la > 0 ? ncart(la - 1) : 0;
    const int rest = ncart(sh[1].l) * ncart(sh[2].l) * ncart(sh[3].l);
    double *Hp = (double *)malloc(sizeof(double) * n0p * rest);
    double *Hm = la > 0 ? (double *)malloc(sizeof(double) * n0m * rest) : NULL;
    eri_cart_callers(&up, &sh[1], &sh[2], &sh[3], kappa, Hp);
    if (la > 0) eri_cart_callers(&dn, &sh[1], &sh[2], &sh[3], kappa, Hm);
    int nout = 1;
    for (int k = 0; k < 4; ++k) nout *= (b->spher && sh[k].l >= 2) ? 2 * sh[k].l + 1 : ncart(sh[k].l);
    double *X = (double *)malloc(sizeof(double) * n0 * rest);
    for (int d = 0; d < 3; ++d) {
        for (int i = 0; i < n0; ++i) {
            int lx, ly, lz;
            cart_comp(la, i, &lx, &ly, &lz);
            const int lq = d == 0 ? lx : (d == 1 ? ly : lz);
            const int ip = cart_idx(lx + (d == 0), ly + (d == 1), lz + (d == 2));
            const int im = lq > 0 ? cart_idx(lx - (d == 0), ly - (d == 1), lz - (d == 2)) : -1;
            for (int r = 0; r < rest; ++r) {
                double v = Hp[ip + n0p * r];
                if (im >= 0) v -= (double)lq * Hm[im + n0m * r];
                X[i + n0 * r] = v;
            }
        }
        int nn[4];
        for (int k = 0; k < 4; ++k) nn[k] = ncart(sh[k].l);
        if (b->spher)
            for (int k = 0; k < 4; ++k)
                if (sh[k].l >= 2) {
                    double U[11 * 66];
                    or_spher_u(sh[k].l, U);
                    transform_index(X, nn, k, U, 2 * sh[k].l + 1);
                }
        double *out = dV + (size_t)d * nout;
        for (int i3 = 0; i3 < nn[3]; ++i3)
            for (int i2 = 0; i2 < nn[2]; ++i2)
                for (int i1 = 0; i1 < nn[1]; ++i1)
                    for (int i0 = 0; i0 < nn[0]; ++i0) {
                        const int q = i0 + nn[0] * (i1 + nn[1] * (i2 + nn[2] * i3));
                        out[q] = X[q] * sh[0].nrm[i0] * sh[1].nrm[i1] * sh[2].nrm[i2] * sh[3].nrm[i3];
                    }
    }
    free(X);
    free(Hp);
    free(Hm);
    return 0;
}
