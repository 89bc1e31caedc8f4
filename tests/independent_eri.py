"""Independent ERI evaluator used only to pin the oracle.

Obara-Saika vertical recurrence for [e0|f0]^(m) on primitive Cartesian
Gaussians (Obara & Saika, JCP 84, 3963 (1986), eqs. for the electron
repulsion integrals) followed by the Head-Gordon-Pople horizontal
recurrence, with Boys values from mpmath at 30 digits (F_m(x) = 1F1(m+1/2;
m+3/2; -x)/(2m+1)).  This is a different recurrence scheme from the
McMurchie-Davidson Hermite expansion the oracle restates, and it uses exact
Boys values instead of the tabulated ones, so agreement pins the oracle's
algebra up to the table's interpolation error.
"""
from __future__ import annotations

import math
from functools import lru_cache

import mpmath
import numpy as np

from gigpu.basis import cart_components, spher_u

mpmath.mp.dps = 30


def boys_exact(mmax, x):
    out = []
    for m in range(mmax + 1):
        if x == 0.0:
            out.append(1.0 / (2 * m + 1))
        else:
            out.append(float(mpmath.hyp1f1(m + 0.5, m + 1.5, -x) / (2 * m + 1)))
    return out


def _prim_eri_cart(la, lb, lc, ld, A, B, C, D, a, b, c, d):
    """Cartesian primitive (ab|cd) for unnormalized x^l exp(-alpha r^2)."""
    A, B, C, D = map(np.asarray, (A, B, C, D))
    p = a + b
    q = c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    W = (p * P + q * Q) / (p + q)
    rho = p * q / (p + q)
    Kab = math.exp(-a * b / p * float(np.dot(A - B, A - B)))
    Kcd = math.exp(-c * d / q * float(np.dot(C - D, C - D)))
    pref = 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd
    L = la + lb + lc + ld
    F = boys_exact(L, rho * float(np.dot(P - Q, P - Q)))
    PA, QC, WP, WQ = P - A, Q - C, W - P, W - Q

    @lru_cache(maxsize=None)
    def vrr(e, f, m):
        # [e 0 | f 0]^(m), e and f Cartesian exponent triples
        if any(v < 0 for v in e) or any(v < 0 for v in f):
            return 0.0
        if sum(e) == 0 and sum(f) == 0:
            return pref * F[m]
        if sum(e) > 0:
            i = 0 if e[0] > 0 else (1 if e[1] > 0 else 2)
            em = list(e); em[i] -= 1; em = tuple(em)
            emm = list(em); emm[i] -= 1; emm = tuple(emm)
            fm_ = list(f); fm_[i] -= 1; fm_ = tuple(fm_)
            val = PA[i] * vrr(em, f, m) + WP[i] * vrr(em, f, m + 1)
            if em[i] > 0:
                val += em[i] / (2 * p) * (vrr(emm, f, m) - rho / p * vrr(emm, f, m + 1))
            if f[i] > 0:
                val += f[i] / (2 * (p + q)) * vrr(em, fm_, m + 1)
            return val
        i = 0 if f[0] > 0 else (1 if f[1] > 0 else 2)
        fm_ = list(f); fm_[i] -= 1; fm_ = tuple(fm_)
        fmm = list(fm_); fmm[i] -= 1; fmm = tuple(fmm)
        val = QC[i] * vrr(e, fm_, m) + WQ[i] * vrr(e, fm_, m + 1)
        if fm_[i] > 0:
            val += fm_[i] / (2 * q) * (vrr(e, fmm, m) - rho / q * vrr(e, fmm, m + 1))
        return val

    AB = A - B
    CD = C - D

    @lru_cache(maxsize=None)
    def hrr(e, bb, f, dd):
        # (e bb | f dd) by transfer onto B then D
        if sum(dd) > 0:
            i = 0 if dd[0] > 0 else (1 if dd[1] > 0 else 2)
            d1 = list(dd); d1[i] -= 1; d1 = tuple(d1)
            f1 = list(f); f1[i] += 1; f1 = tuple(f1)
            return hrr(e, bb, f1, d1) + CD[i] * hrr(e, bb, f, d1)
        if sum(bb) > 0:
            i = 0 if bb[0] > 0 else (1 if bb[1] > 0 else 2)
            b1 = list(bb); b1[i] -= 1; b1 = tuple(b1)
            e1 = list(e); e1[i] += 1; e1 = tuple(e1)
            return hrr(e1, b1, f, dd) + AB[i] * hrr(e, b1, f, dd)
        return vrr(e, f, 0)

    ca, cb, cc, cd = (cart_components(l) for l in (la, lb, lc, ld))
    out = np.zeros((len(ca), len(cb), len(cc), len(cd)))
    for i, ea in enumerate(ca):
        for j, eb in enumerate(cb):
            for k, ec in enumerate(cc):
                for l_, ed in enumerate(cd):
                    out[i, j, k, l_] = hrr(ea, eb, ec, ed)
    return out


def shell_quartet(basis, shells):
    """Normalized (spherical or Cartesian) block V[a,b,c,d] for 0-based shells."""
    data = []
    for s in shells:
        p = basis.shell_params_idx[s] - 1
        l = int(basis.shell_momentum[p])
        n = int(basis.nprimitives[p])
        data.append((l, basis.atom_coords[basis.shell_centers[s] - 1], basis.exponents[p, :n],
                     basis.cntr_coeffs[p, :n], basis.norm_factors[p]))
    (la, A, ea, ca, na), (lb, B, eb, cb, nb), (lc, Cc, ec, cc, nc), (ld, D, ed, cd, nd) = data
    V = 0.0
    for i in range(len(ea)):
        for j in range(len(eb)):
            for k in range(len(ec)):
                for m in range(len(ed)):
                    V = V + ca[i] * cb[j] * cc[k] * cd[m] * _prim_eri_cart(
                        la, lb, lc, ld, A, B, Cc, D, ea[i], eb[j], ec[k], ed[m])
    ls = [la, lb, lc, ld]
    if basis.spher_ao:
        for ax, l in enumerate(ls):
            if l >= 2:
                U = spher_u(l)
                V = np.moveaxis(np.tensordot(U, V, axes=([1], [ax])), 0, ax)
    norms = [na, nb, nc, nd]
    for ax in range(4):
        nrm = norms[ax][:V.shape[ax]]
        shape = [1, 1, 1, 1]
        shape[ax] = -1
        V = V * nrm.reshape(shape)
    return V
