"""Dense restatement of Step 1 of the two-step Cholesky decomposition, written
from the reference's Fortran (not from the oracle's C), for tiny bases:

  chol2_Step1_FullInterface    TwoStepCholesky.f90:119-160 (MaxNCholesky,
                               MaxNAngFunc)
  chol2_Vdiag / write_Vdiag    TwoStepCholesky_Step1.f90:175-279 (shell pairs
                               ShB = 1..N, ShA = ShB..N; b outer, a inner,
                               a >= b for ShA = ShB)
  chol2_DefineBase             TwoStepCholesky_Step1.f90:282-442 (ascending
                               DSORT0 of the pair maxima, loop to the first
                               pair above min(tau, 1e-8)^2 / MaxDrs; the last
                               discarded pair, k0, stays)
  chol2_DefineCandidates       TwoStepCholesky_Step1.f90:533-575
  chol2_DefineBatch            TwoStepCholesky_Step1.f90:578-625
  chol2_SubtractR / chol2_M    TwoStepCholesky_Step1.f90:628-691,733-856
  chol2_BatchMax               TwoStepCholesky_Step1.f90:694-730 (strict >,
                               batch order)
  chol2_Lj                     TwoStepCholesky_Step1.f90:859-881
  chol2_Pivots                 TwoStepCholesky_Step1.f90:884-1103 (loop and
                               stop rules)

The Coulomb matrix is a dense array over every orbital pair, so M is a slice
of V minus the earlier vectors, no compressed storage is kept and no
quartet is evaluated twice; the heuristics (sorts, thresholds, batch limit,
first maximum) are applied exactly as the Fortran applies them.  DSORT0 is
the label-by-label transcription in test_oracle_sort.py.

Comparison keys: with quantized=False the raw doubles, as the reference
compares them.  With quantized=True the canonicalization the library and the
oracle document for near-ties: pair maxima of the BASE sort rounded to a
26-bit mantissa; residual diagonals in the candidate and batch sorts and in
the batch maximum rounded to multiples of delta = min(2^(floor(log2 MaxDrs) -
32), 2^(floor(log2 tau) - 8)); thresholds always compare raw values."""
import math

import numpy as np

from test_oracle_sort import dsort0_labels

SIGMA = 0.01        # spread factor of Dmin (TwoStepCholesky_Step1.f90:1004)
MIN_BATCH = 1000    # MaxNBatch = min(NCand, max(MaxNAngFunc^2, 1000)) (:998)
KEEP_K0 = True      # DefineBase keeps the last discarded shell pair (:391-419)


def vdiag_shell_pairs(nshells):
    """(ShA, ShB), 1-based, in the order of chol2_Vdiag."""
    return [(a, b) for b in range(1, nshells + 1) for a in range(b, nshells + 1)]


def _qrel(x):
    if x == 0.0 or not math.isfinite(x):
        return x
    m, e = math.frexp(x)
    return math.ldexp(round(math.ldexp(m, 27)), e - 27)   # round half to even, like rint


def step1(V, basis, tau, quantized=True):
    """Pivots of Step 1.  V: dense (pq|rs) over the orbital pairs of
    vdiag_shell_pairs(basis.nshells) in that order.  Returns (pivot_order,
    shell_pairs, nvecs, max_residual): pivot_order as 1-based FULL storage
    locations in BASE order (TChol2Vecs Pivots numbering), shell_pairs the
    BASE ShellPairs (1-based)."""
    nsh = basis.nshells
    nang = [basis.shell_nang(s) for s in range(nsh)]
    lmax = max(basis.shell_l(s) for s in range(nsh))
    max_nang = 2 * lmax + 1 if basis.spher_ao else (lmax + 1) * (lmax + 2) // 2
    sp_all = vdiag_shell_pairs(nsh)
    dims, locs = [], []
    loc = 0
    for a, b in sp_all:
        n = nang[a - 1] * nang[b - 1] if a != b else nang[a - 1] * (nang[a - 1] + 1) // 2
        dims.append(n)
        locs.append(loc)
        loc += n
    assert V.shape == (loc, loc)
    vdiag = np.diag(V).copy()
    # ---- chol2_DefineBase
    vmax = [float(vdiag[locs[k]:locs[k] + dims[k]].max()) for k in range(len(sp_all))]
    maxdrs = max(vmax)
    assert maxdrs > 0.0
    pthresh = min(tau, 1.0e-8) ** 2 / maxdrs
    keys = [_qrel(v) for v in vmax] if quantized else vmax
    _, idx = dsort0_labels(keys, list(range(1, len(sp_all) + 1)), 2)
    k0 = 1
    for k, l in enumerate(idx, start=1):
        if vmax[l - 1] <= pthresh:
            k0 = k
        else:
            break
    first = k0 - 1
    if not KEEP_K0 and vmax[idx[first] - 1] <= pthresh:
        first += 1
    base = [l - 1 for l in idx[first:]]                  # Vdiag-order pair numbers, BASE order
    nsp = len(base)
    rows_of = []                                          # FULL rows of each BASE pair into V
    full_loc = []
    p = 0
    for s in base:
        rows_of.append(np.arange(locs[s], locs[s] + dims[s]))
        full_loc.append(p)
        p += dims[s]
    nop = p
    Vb = V[np.concatenate(rows_of)][:, np.concatenate(rows_of)]   # V over BASE, FULL order
    base_rows = [np.arange(full_loc[i], full_loc[i] + dims[base[i]]) for i in range(nsp)]
    D = np.diag(Vb).copy()
    max_nchol = min(math.ceil(basis.nao * 15.0), nop)
    # ---- residual key quantum
    _, e_m = math.frexp(maxdrs)
    _, e_t = math.frexp(tau)
    inv_delta = math.ldexp(1.0, max(33 - e_m, 9 - e_t))

    def key(x):
        return float(np.rint(x * inv_delta)) if quantized else float(x)

    # ---- chol2_Pivots
    cand = list(range(nsp))
    L = np.zeros((nop, 0))
    pivot_order = []
    nchol = 0
    dmax = float(D.max())
    converged = False
    do_macro = max_nchol > 0
    while do_macro:
        # chol2_DefineCandidates
        raw = {s: float(D[base_rows[s]].max()) for s in cand}
        _, order = dsort0_labels([key(raw[s]) for s in cand], [s + 1 for s in cand], -2)
        order = [s - 1 for s in order]
        new = []
        for s in order:
            if raw[s] > tau:
                new.append(s)
            elif not quantized:
                break
        cand = new
        crow = np.concatenate([base_rows[s] for s in cand]) if cand else np.zeros(0, int)
        ncand = len(crow)
        max_nbatch = min(ncand, max(max_nang ** 2, MIN_BATCH))
        dmin = max(SIGMA * dmax, tau)
        # chol2_DefineBatch
        _, border = dsort0_labels([key(raw[s]) for s in cand], [s + 1 for s in cand], -2)
        batch = []
        nb = 0
        for s in (b - 1 for b in border):
            d = dims[base[s]]
            if raw[s] > dmin and nb + d <= max_nbatch:
                batch.append(s)
                nb += d
            else:
                break
        bcol = np.concatenate([base_rows[s] for s in batch]) if batch else np.zeros(0, int)
        # chol2_M + chol2_SubtractR: columns of the residual matrix
        M = Vb[np.ix_(crow, bcol)] - L[crow] @ L[bcol].T
        pos = {r: i for i, r in enumerate(crow)}

        def batch_max():
            best_k, best_v, best_c = key(-1.0), -1.0, -1
            for c, r in enumerate(bcol):
                k = key(D[r])
                if k > best_k:
                    best_k, best_v, best_c = k, float(D[r]), c
            return best_v, best_c

        maxdpq, c = batch_max()
        j = 0
        Lm = np.zeros((ncand, 0))
        do_micro = nchol < max_nchol and nb > 0
        while do_micro:
            j += 1
            nchol += 1
            r = bcol[c]
            pivot_order.append(r + 1)
            # chol2_Lj
            l = M[:, c] - Lm @ Lm[pos[r], :]
            l = l / math.sqrt(maxdpq)
            Lm = np.column_stack([Lm, l])
            D[crow] -= l ** 2
            maxdpq, c = batch_max()
            if maxdpq <= dmin:
                do_micro = False
            if j == nb or nchol == max_nchol:
                do_micro = False
        if j > 0:
            Lnew = np.zeros((nop, j))
            Lnew[crow] = Lm
            L = np.column_stack([L, Lnew])
        dmax = float(D[crow].max()) if ncand else -math.inf
        if dmax <= tau:
            converged = True
            do_macro = False
        if nchol == max_nchol or nb == 0:
            do_macro = False
    shell_pairs = np.array([sp_all[s] for s in base], dtype=np.int64)
    return np.array(pivot_order, dtype=np.int64), shell_pairs, nchol, dmax, converged
