"""Benchmark of the Cholesky/ERI hot path on one MI355X.

One step = one full two-step Cholesky decomposition of the Coulomb matrix of
the (H2O)16 water cluster (10 A box, O 8s5p3d1f / H 4s2p1d, 1200 spherical
AOs) at tau = 1e-6, from basis upload to Rkpq resident in HBM, followed by
the AO->MO transform of the vectors (random orthonormal C, nocc = 40,
nvirt = 300, CA/CB/Rkab in device memory).

Prints one JSON line.  value = ERI shell quartets evaluated per second of
Cholesky build wall time (whole job, one GPU); cholesky_build_s is the
build wall time itself.  The roofline object covers the ERI class kernels
(the dominant kernel family): algorithmic FP64 operations (per-class counts
emitted by gigpu/eri_gen.py times quartets) over their HIP-event time on the
library stream.  cpu_baseline times the CPU oracle (the plain-C restatement
of the reference algorithm, OpenMP) on a bounded sample: the full build of
the (H2O)5 cluster with the same basis (about 10 s on 16 host cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

# Hardware queues per process (HIP's default is 4): the batched form runs one
# stream per molecule, and with 4 queues at most four kernels run
# concurrently (measured: 0.60 s with 4, 0.55 s with 16).  Read by the HIP
# runtime at initialisation, before any HIP call; set here even when the
# environment already holds a value.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gammcor-integrals_amd"))

METRIC = "Cholesky build s + ERI quartets/s, batched/mixed vs plain GPU at tau=1e-6"
FP64_PEAK_TFLOPS = 78.6


class Hip:
    """Minimal device-memory plumbing for the device-pointer MO transform."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        if self.L.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        return p.value

    def h2d(self, dst, arr):
        if self.L.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy failed")

    def free(self, p):
        self.L.hipFree(C.c_void_p(p))

    def sync(self):
        self.L.hipDeviceSynchronize()


def profiled_traffic():
    """HBM bytes per ERI launch measured by the committed counter profile of
    this workload, or None."""
    path = os.path.join(ROOT, "profiles", "eri_kernels_h2o16_hbm_bytes.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def batched_form(nworkers: int, autotune_batch: bool = True):
    """The batched form on 64 independent molecules ((H2O)n, n uniform in
    1..12, O 6s4p2d / H 3s1p): the same decompositions one at a time (plain)
    and concurrently (gi_chol_build_batch), after one warm-up pass; then the
    autotuner's settings for this workload (tile sizes, lane kernels, vector
    blocks, workers) and the batched time with them."""
    from gigpu import lib, systems
    mols = systems.batch_systems(64, seed=3)
    ctxs = [lib.Context(0) for _ in mols]
    try:
        for c, b in zip(ctxs, mols):
            c.set_basis(b)
        lib.chol_build_batch(ctxs, tau=1e-6, nworkers=nworkers)
        t0 = time.perf_counter()
        for c in ctxs:
            c.chol_build(tau=1e-6)
        t_plain = time.perf_counter() - t0
        nv_plain = [c.chol_dims()["nvecs"] for c in ctxs]
        quartets = sum(c.chol_dims()["nquartets"] for c in ctxs)
        t0 = time.perf_counter()
        lib.chol_build_batch(ctxs, tau=1e-6, nworkers=nworkers)
        t_batch = time.perf_counter() - t0
        nv_batch = [c.chol_dims()["nvecs"] for c in ctxs]
        tuned = None
        if autotune_batch:
            from gigpu import autotune
            best, _, trials = autotune.tune(ctxs, nworkers=[nworkers, 8, 16], reps=1)
            ts = []
            for _ in range(2):
                t0 = time.perf_counter()
                lib.chol_build_batch(ctxs, tau=1e-6, nworkers=best["nworkers"])
                ts.append(time.perf_counter() - t0)
            tuned = {"settings": best, "batched_s": float(min(ts)), "trials": len(trials),
                     "same_nvecs": [c.chol_dims()["nvecs"] for c in ctxs] == nv_plain}
    finally:
        for c in ctxs:
            c.close()
    return {"workload": "64 independent (H2O)n, n uniform in 1..12 (seed 3), O 6s4p2d / H 3s1p, tau=1e-6",
            "nao_total": int(sum(b.nao for b in mols)), "nvecs_total": int(sum(nv_batch)),
            "same_nvecs": nv_plain == nv_batch, "quartets": int(quartets), "workers": nworkers,
            "plain_s": t_plain, "batched_s": t_batch, "speedup": t_plain / t_batch,
            "plain_quartets_per_s": quartets / t_plain, "batched_quartets_per_s": quartets / t_batch,
            "autotuned": tuned}


def precision_form(ctx, basis, Cmo, hip, reps: int = 3):
    """Split-BF16 MFMA products (four 8-bit slices per element, exact FP32
    sums per 256 k, FP64 across chunks) against the FP64 path on the
    same (H2O)16 system (1200 AOs, the nbf of the precision configuration):
    the AO->MO transform with nocc=80, nvirt=600 (device timings, max error
    over all Rkab elements) and the Step-1 trailing update (build time,
    vectors, and the reconstruction error max_pq (pq|pq) - sum_k R_k(pq)^2,
    both with the Step-1 pivot decode so that the FP64 figure is the bound
    tau)."""
    from gigpu import lib
    out = {}
    ctx.set_basis(basis)
    ctx.chol_build(tau=1e-6)
    ref_res = ctx.chol_residual()[0]
    nocc, nvirt = 80, 600
    nv = ctx.chol_dims()["nvecs"]
    dC = hip.alloc(Cmo.nbytes)
    hip.h2d(dC, Cmo)
    dR = hip.alloc(8 * nv * nocc * nvirt)
    times = {}
    for name, fl in (("fp64", 0), ("split_bf16", lib.MO_SPLIT_BF16)):
        ctx.mo_transform_device(dC, basis.nao, 1, nocc, dC, basis.nao, nocc + 1, nocc + nvirt, dR, flags=fl)
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.mo_transform_device(dC, basis.nao, 1, nocc, dC, basis.nao, nocc + 1, nocc + nvirt, dR, flags=fl)
            t.append(time.perf_counter() - t0)
        times[name] = float(np.median(t))
    hip.free(dR)
    hip.free(dC)
    r64 = ctx.mo_transform(Cmo, 1, nocc, Cmo, nocc + 1, nocc + nvirt)
    r16 = ctx.mo_transform(Cmo, 1, nocc, Cmo, nocc + 1, nocc + nvirt, flags=lib.MO_SPLIT_BF16)
    err = float(np.abs(r16 - r64).max())
    rmax = float(np.abs(r64).max())
    del r16, r64
    out["mo_transform"] = {"nocc": nocc, "nvirt": nvirt, "nvecs": nv, "fp64_s": times["fp64"],
                           "split_bf16_s": times["split_bf16"], "speedup": times["fp64"] / times["split_bf16"],
                           "max_abs_err": err, "max_rel_err": err / rmax}
    tu = {}
    for name, fl in (("fp64", lib.CHOL_DECODE_FIXED), ("split_bf16", lib.CHOL_DECODE_FIXED | lib.CHOL_SPLIT_BF16)):
        ctx.chol_build(tau=1e-6, flags=fl)
        t0 = time.perf_counter()
        ctx.chol_build(tau=1e-6, flags=fl)
        t_b = time.perf_counter() - t0
        tm = ctx.chol_timings()
        mx, mn = ctx.chol_residual()
        tu[name] = {"build_s": t_b, "subtract_r_s": tm["rtr"], "nvecs": ctx.chol_dims()["nvecs"],
                    "residual_max": mx, "residual_min": mn}
    out["trailing_update"] = tu
    out["residual_max_reference_decode"] = ref_res
    return out


def gradient_form(reps: int = 3, large: bool = True):
    """The differentiation form on (H2O)8, O 6s4p2d / H 3s1p (320 AOs, 72
    coordinates), D = C_occ C_occ^T with nocc = 40: median seconds of the
    energy-only call and of the energy + nuclear gradient call, beside what
    central differences over all coordinates would cost (2 energies each),
    and what density-weighted Schwarz screening at 1e-12 changes; then one
    energy and one energy + gradient call on the 1200-AO (H2O)16 system of
    the headline, screened at 1e-12."""
    from gigpu import lib, systems
    basis = systems.config_gradient_cluster()
    D = systems.occupied_density(basis.nao, 40, seed=7)
    ctx = lib.Context(0)
    try:
        ctx.set_basis(basis)
        ctx.coulomb_energy(D)
        te, tg = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            e = ctx.coulomb_energy(D, grad=False)
            te.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            e2, g = ctx.coulomb_energy(D)
            tg.append(time.perf_counter() - t0)
        # density-weighted Schwarz screening at 10^-12 on the same system:
        # its effect on the result
        ctx.set_option("grad_screen", 12)
        es, gs = ctx.coulomb_energy(D)
        ctx.set_option("grad_screen", 0)
    finally:
        ctx.close()
    nsh = basis.nshells
    npair = nsh * (nsh + 1) // 2
    t_e, t_g = float(np.median(te)), float(np.median(tg))
    out = {"workload": "E_J = 1/2 sum D_pq (pq|rs) D_rs and dE_J/dR of (H2O)8 (seed 5), O 6s4p2d / H 3s1p, "
                       "D = C_occ C_occ^T (nocc=40, seed 7)",
           "nao": basis.nao, "natoms": basis.natoms, "shell_pairs": npair,
           "pair_products": npair * npair, "energy": e, "energy_s": t_e, "gradient_s": t_g,
           "gradient_over_energy": t_g / t_e, "central_differences_s": 2 * 3 * basis.natoms * t_e,
           "max_abs_gradient": float(np.abs(g).max()),
           "screened_1e-12": {"energy_diff": float(es - e2), "max_abs_gradient_diff": float(np.abs(gs - g).max())}}
    if large:
        # the (H2O)16 system of the headline (1200 AOs, O 8s5p3d1f / H 4s2p1d),
        # screened at 10^-12 (unscreened: 22 s energy, 26 s energy + gradient)
        lb = systems.water_system(16, 10.0, "large", seed=2)
        LD = systems.occupied_density(lb.nao, 40, seed=7)
        ctx = lib.Context(0)
        try:
            ctx.set_basis(lb)
            ctx.set_option("grad_screen", 12)
            t0 = time.perf_counter()
            le = ctx.coulomb_energy(LD, grad=False)
            lte = time.perf_counter() - t0
            t0 = time.perf_counter()
            le2, lg = ctx.coulomb_energy(LD)
            ltg = time.perf_counter() - t0
        finally:
            ctx.close()
        out["h2o16"] = {"workload": "the same quantities for (H2O)16 (seed 2), O 8s5p3d1f / H 4s2p1d, 1200 AOs, "
                                    "D = C_occ C_occ^T (nocc=40, seed 7), Schwarz screening 1e-12",
                        "nao": lb.nao, "natoms": lb.natoms, "energy": le2, "energy_s": lte, "gradient_s": ltg,
                        "max_abs_gradient": float(np.abs(lg).max())}
    return out


def contracted_form(reps: int = 2):
    """The headline build on contracted shells: (H2O)16 in the same box with
    the same shells per element and l (1200 AOs), the first of them contracted
    from three primitives and the second from two (2560 primitive functions
    in all), tau = 1e-6; best of reps builds, ERI kernels timed with HIP
    events."""
    from gigpu import lib, systems
    basis = systems.water_system(16, 10.0, "large", seed=2, contracted=True)
    ctx = lib.Context(0)
    try:
        ctx.set_basis(basis)
        ctx.chol_build(tau=1e-6, flags=lib.CHOL_TIME_ERI)
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.chol_build(tau=1e-6, flags=lib.CHOL_TIME_ERI)
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, ctx.chol_dims(), ctx.chol_timings())
    finally:
        ctx.close()
    t, d, tm = best
    nprim = int(sum(basis.nprimitives[p - 1] * (2 * basis.shell_momentum[p - 1] + 1)
                    for p in basis.shell_params_idx))
    return {"workload": "(H2O)16 (seed 2), 10 A box, O 8s5p3d1f / H 4s2p1d segmented contractions "
                        "(3 and 2 primitives in the two tightest shells of each l), tau=1e-6",
            "nao": basis.nao, "primitive_functions": nprim, "nvecs": d["nvecs"], "build_s": t,
            "eri_quartets": d["nquartets"], "quartets_per_s": d["nquartets"] / t,
            "eri_kernel_s": tm["eri_kernel"], "eri_tflops": tm["eri_flops"] / tm["eri_kernel"] / 1e12
            if tm["eri_kernel"] > 0 else None}


def thc_flops(npts, nvec, nao, npairs):
    """Algorithmic FP64 operations of one least-squares Z: XR over the npairs
    stored orbital pairs of the decomposition (the pairs that survive its
    screening), S = (X X^T)^2, the Cholesky factor of S and the two
    triangular solves with nvec right-hand sides."""
    xr = 2.0 * npts * nvec * npairs
    s = 2.0 * npts * npts * nao
    solve = npts ** 3 / 3.0 + 2.0 * npts * npts * nvec
    return xr, s, solve


# radii (bohr) of the synthetic parent grid of the pruning timing: 8 radii x
# 50 directions per atom
PARENT_RADII = (0.2, 0.45, 0.8, 1.2, 1.7, 2.4, 3.3, 4.5)


def thc_form(hip, reps: int = 3, cpu: bool = True):
    """Tensor hypercontraction on the 1200-AO (H2O)16 system of the headline:
    the decomposition at the THC default threshold (tau = 1e-5), 3072
    atom-centred grid points (48 atoms x 4 radii x 16 directions), then the
    orbitals on the grid, the least-squares Z and Rkab for nocc = 40 x
    nvirt = 300, all on device pointers; median of reps calls each; then the
    pruning of a 19,200-point parent grid (48 atoms x 8 radii x 50
    directions) at QRThresh 1e-3.  The CPU leg times the oracle's Z (NumPy,
    threaded BLAS, LAPACK dsysv) and its pruning on a bounded sample, (H2O)4
    with the same basis and grid rules (768 points; 2880 parent points, 30
    directions), beside the device on the same sample."""
    from gigpu import lib, systems

    def one(basis, npts_dir, reps_):
        pts = np.ascontiguousarray(systems.atom_centred_points(basis, ndir=npts_dir, seed=21))
        n, nao = len(pts), basis.nao
        ctx = lib.Context(0)
        bufs = []
        try:
            ctx.set_basis(basis)
            ctx.chol_build(tau=1e-5)
            dims = ctx.chol_dims()
            nv = dims["nvecs"]
            dxyz = hip.alloc(pts.nbytes)
            hip.h2d(dxyz, pts)
            dX, dZ = hip.alloc(8 * n * nao), hip.alloc(8 * n * nv)
            bufs += [dxyz, dX, dZ]
            t_orb, t_z = [], []
            for _ in range(reps_ + 1):
                t0 = time.perf_counter()
                ctx.thc_orbitals_device(n, dxyz, dX, n)
                t1 = time.perf_counter()
                ctx.thc_z_device(n, dX, n, dZ, n)
                t2 = time.perf_counter()
                t_orb.append(t1 - t0)
                t_z.append(t2 - t1)
            res = {"nao": nao, "npts": n, "nvecs": nv, "norb_pairs": dims["norb_pairs"],
                   "orbitals_s": float(np.median(t_orb[1:])),
                   "z_s": float(np.median(t_z[1:]))}
            return ctx, res, pts, bufs
        except Exception:
            ctx.close()
            for b in bufs:
                hip.free(b)
            raise

    basis = systems.water_system(16, 10.0, "large", seed=2)
    ctx, res, pts, bufs = one(basis, 16, reps)
    try:
        n, nao, nv = res["npts"], res["nao"], res["nvecs"]
        xr, sf, sv = thc_flops(n, nv, nao, res["norb_pairs"])
        res["z_tflops"] = (xr + sf + sv) / res["z_s"] / 1e12
        res["z_frac_fp64_peak"] = res["z_tflops"] / FP64_PEAK_TFLOPS
        res["z_flops"] = {"xr": xr, "metric": sf, "solve": sv}
        nocc, nvirt = 40, 300
        rng = np.random.default_rng(2)
        Cmo = np.asfortranarray(np.linalg.qr(rng.normal(size=(nao, nocc + nvirt)))[0])
        dC, dR = hip.alloc(Cmo.nbytes), hip.alloc(8 * nv * nocc * nvirt)
        bufs += [dC, dR]
        hip.h2d(dC, Cmo)
        dX, dZ = bufs[1], bufs[2]
        tr = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            ctx.thc_rkab_device(n, dX, n, dZ, n, nv, dC, nao, 1, nocc, dC, nao, nocc + 1, nocc + nvirt, dR)
            tr.append(time.perf_counter() - t0)
        res["rkab_s"] = float(np.median(tr[1:]))
        rk_flops = 2.0 * n * nao * (nocc + nvirt) + 2.0 * nv * n * nocc * nvirt
        res["rkab_tflops"] = rk_flops / res["rkab_s"] / 1e12
        # grid pruning (thc_Chol_Pivots) of a parent grid at the default
        # QRThresh = QRThreshReduced = 1e-3, BlockDim = 500
        parent = systems.atom_centred_points(basis, radii=PARENT_RADII, ndir=50, seed=21)
        ctx.thc_grid_pivots(parent[:1000], 1e-3, 1e-3, 500)
        t0 = time.perf_counter()
        piv, nred = ctx.thc_grid_pivots(parent, 1e-3, 1e-3, 500)
        res["pruning"] = {"parent_points": len(parent), "ngrid": len(piv), "ngrid_reduced": nred,
                          "s": time.perf_counter() - t0}
    finally:
        ctx.close()
        for b in bufs:
            hip.free(b)
    out = {"workload": "THC of (H2O)16 (seed 2), O 8s5p3d1f / H 4s2p1d, 1200 AOs, tau=1e-5, 3072 atom-centred "
                       "points (seed 21), Rkab nocc=40 x nvirt=300", **res}
    if cpu:
        # the CPU leg: the oracle's NumPy restatement, timed as the baseline
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import thc as thc_oracle
        small = systems.water_system(4, 6.0, "large", seed=2)
        sctx, sres, spts, sbufs = one(small, 16, reps)
        try:
            m = sctx.chol_meta()
            rk = sctx.chol_rkpq()
            X = sctx.thc_orbitals(spts)
        finally:
            sctx.close()
            for b in sbufs:
                hip.free(b)
        cols = thc_oracle.storage_pairs(small, m["shell_pairs"], m["shell_pair_loc"], m["max_subset_dim"],
                                        m["nsubsets"])
        t0 = time.perf_counter()
        thc_oracle.z(X, cols, rk)
        t_cpu = time.perf_counter() - t0
        sparent = systems.atom_centred_points(small, radii=PARENT_RADII, ndir=30, seed=21)
        t0 = time.perf_counter()
        opiv, _ = thc_oracle.grid_pivots(small, sparent, 1e-3, 1e-3, 500)
        t_cpu_prune = time.perf_counter() - t0
        pctx = lib.Context(0)
        try:
            pctx.set_basis(small)
            pctx.thc_grid_pivots(sparent[:500], 1e-3, 1e-3, 500)
            t0 = time.perf_counter()
            gpiv, _ = pctx.thc_grid_pivots(sparent, 1e-3, 1e-3, 500)
            t_gpu_prune = time.perf_counter() - t0
        finally:
            pctx.close()
        xr, sf, sv = thc_flops(sres["npts"], sres["nvecs"], sres["nao"], sres["norb_pairs"])
        out["cpu_sample"] = {"workload": "the oracle's Z (NumPy + threaded BLAS + LAPACK dsysv) for (H2O)4 (seed 2), "
                                         "same basis, %d points" % sres["npts"],
                             "cores": os.cpu_count() if "OMP_NUM_THREADS" not in os.environ
                             else int(os.environ["OMP_NUM_THREADS"]),
                             "nao": sres["nao"], "nvecs": sres["nvecs"], "cpu_z_s": t_cpu,
                             "gpu_z_s": sres["z_s"], "cpu_tflops": (xr + sf + sv) / t_cpu / 1e12,
                             "gpu_over_cpu": t_cpu / sres["z_s"],
                             "pruning": {"parent_points": len(sparent), "ngrid": len(gpiv), "cpu_s": t_cpu_prune,
                                         "gpu_s": t_gpu_prune, "same_pivots": bool(np.array_equal(opiv, gpiv))}}
    return out


def cpu_baseline(seed: int):
    """The CPU restatement (oracle, C/OpenMP) on a bounded sample -- the full
    build of (H2O)5 with the headline basis -- and, beside it, the GPU build
    of the same system, so that the two rates are on one workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from gigpu import lib, systems
    pyoracle.build()
    b = systems.water_system(5, 10.0, "large", seed=seed)
    o = pyoracle.Oracle(b)
    t0 = time.perf_counter()
    r = o.chol(tau=1e-6)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ctx = lib.Context(0)
    try:
        ctx.set_basis(b)
        ctx.chol_build(tau=1e-6)
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            ctx.set_basis(b)
            ctx.chol_build(tau=1e-6)
            ts.append(time.perf_counter() - t1)
        gd = ctx.chol_dims()
        same_pivots = bool(np.array_equal(ctx.chol_meta()["pivot_order"], r.pivot_order))
    finally:
        ctx.close()
    tg = float(np.median(ts))
    return {"value": r.nquartets / dt, "unit": "quartets/s", "cores": threads, "kind": "port",
            "sample": "CPU restatement (oracle, C/OpenMP), full two-step Cholesky of (H2O)5 with the same "
                      "basis (%d AOs, %d vectors, %d quartets) in %.2f s" % (b.nao, r.nvecs, r.nquartets, dt),
            "gpu_same_sample": {"build_s": tg, "quartets_per_s": gd["nquartets"] / tg, "nvecs": gd["nvecs"],
                                "same_pivot_order": same_pivots, "speedup_vs_cpu": dt / tg}}


def decode_and_ties(ctx, basis, ref_order):
    """Beside the headline build (the reference's chol2_pq2p_ge_q decode of
    same-shell pivots in Step 2, reproduced for parity): the build with the
    Step-1 decode (GI_CHOL_DECODE_FIXED), whose reconstruction error is the
    bound tau, and the build with raw (unquantized) comparison keys
    (GI_CHOL_RAW_TIES), compared pivot by pivot with the headline build."""
    from gigpu import lib
    out = {}
    ctx.set_basis(basis)
    ctx.chol_build(tau=1e-6, flags=lib.CHOL_DECODE_FIXED)
    t0 = time.perf_counter()
    ctx.set_basis(basis)
    ctx.chol_build(tau=1e-6, flags=lib.CHOL_DECODE_FIXED)
    tb = time.perf_counter() - t0
    mx, mn = ctx.chol_residual()
    out["fixed_decode"] = {"build_s": tb, "nvecs": ctx.chol_dims()["nvecs"], "residual_max": mx,
                           "residual_min": mn}
    ctx.chol_build(tau=1e-6, flags=lib.CHOL_RAW_TIES | lib.CHOL_SKIP_VECTORS)
    raw = ctx.chol_meta()["pivot_order"]
    n = min(len(raw), len(ref_order))
    out["raw_ties"] = {"nvecs": int(len(raw)), "same_pivot_order": bool(np.array_equal(raw, ref_order)),
                       "first_difference": (int(np.argmax(raw[:n] != ref_order[:n]))
                                            if np.any(raw[:n] != ref_order[:n]) else None),
                       "positions_differing": int(np.count_nonzero(raw[:n] != ref_order[:n]) + abs(len(raw) - len(ref_order)))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nwater", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mo", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-form comparison")
    ap.add_argument("--batch-workers", type=int, default=16)
    ap.add_argument("--no-autotune", action="store_true", help="skip autotuning the batched form")
    ap.add_argument("--no-gradient", action="store_true", help="skip the differentiation-form timing")
    ap.add_argument("--no-thc", action="store_true", help="skip the tensor-hypercontraction timing")
    ap.add_argument("--no-contracted", action="store_true", help="skip the contracted-basis build")
    ap.add_argument("--no-precision", action="store_true", help="skip the split-BF16 precision-form comparison")
    ap.add_argument("--no-decode-ties", action="store_true",
                    help="skip the fixed-decode and raw-tie builds beside the headline build")
    ap.add_argument("--eri-stats", action="store_true",
                    help="print per-class ERI kernel statistics to stderr (sets GIGPU_ERI_STATS)")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="launch-layer option of the headline context (gi_set_option), e.g. eri_md4=0")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before the run, one device-to-device copy of this many MiB (a dispatch of known "
                         "size for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (Step 2's vectors would shard by "
              "subset; the sequential Step-1 pivot search stays on one GPU)",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return

    if args.eri_stats:
        os.environ["GIGPU_ERI_STATS"] = "1"
    from gigpu import lib, systems
    basis = systems.water_system(args.nwater, 10.0, "large", seed=args.seed)
    nocc, nvirt = 40, 300
    rng = np.random.default_rng(args.seed)
    Cmo, _ = np.linalg.qr(rng.normal(size=(basis.nao, basis.nao)))
    Cmo = np.asfortranarray(Cmo)
    ctx = lib.Context(0)
    for opt in args.option:
        name, value = opt.split("=", 1)
        ctx.set_option(name, int(value))
    hip = Hip()
    if args.copy_probe_mb > 0:
        nbytes = args.copy_probe_mb << 20
        src, dst = hip.alloc(nbytes), hip.alloc(nbytes)
        if hip.L.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, 3) != 0:
            raise RuntimeError("hipMemcpy failed")
        hip.sync()
        hip.free(src)
        hip.free(dst)
    dC = hip.alloc(Cmo.nbytes)
    hip.h2d(dC, Cmo)
    dR = None

    def step():
        nonlocal dR
        t0 = time.perf_counter()
        ctx.set_basis(basis)
        ctx.chol_build(tau=1e-6, flags=lib.CHOL_TIME_ERI)
        t_build = time.perf_counter() - t0
        d = ctx.chol_dims()
        t_mo = 0.0
        if not args.no_mo:
            if dR is None:
                dR = hip.alloc(8 * d["nvecs"] * nocc * nvirt)
            t1 = time.perf_counter()
            ctx.mo_transform_device(dC, basis.nao, 1, nocc, dC, basis.nao, nocc + 1, nocc + nvirt, dR)
            t_mo = time.perf_counter() - t1
        return t_build, t_mo, d, ctx.chol_timings()

    for _ in range(args.warmup):
        step()
    hip.sync()
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(step())
    hip.sync()
    wall = time.perf_counter() - t0

    t_build = float(np.mean([r[0] for r in results]))
    t_mo = float(np.mean([r[1] for r in results]))
    d = results[-1][2]
    tm = results[-1][3]
    quartets = d["nquartets"]
    eri_s = tm["eri_kernel"]
    eri_flops = tm["eri_flops"]
    launches = int(tm["eri_launches"])
    # algorithmic FP64 ops per launch / average launch duration (= total / total)
    achieved = eri_flops / eri_s / 1e12 if eri_s > 0 else None
    # Rkpq = Inv_L Vabrs^T with Inv_L lower triangular, one GEMM per block row
    # of a multiple of 128 rows counted from the last row (the library's
    # partition for the default vec_blocks = 20)
    nv = d["nvecs"]
    nblk = max(1, min(20, nv // 256))
    rows_blk = max(128, (nv + 64 * nblk) // (128 * nblk) * 128)
    gemm_flops, r1 = 0.0, nv
    while r1 > 0:
        r0 = max(0, r1 - rows_blk)
        gemm_flops += 2.0 * (r1 - r0) * r1
        r1 = r0
    gemm_flops *= d["norb_pairs"]
    out = {
        "metric": METRIC,
        "value": quartets / t_build,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded (H2O)%d cluster, even-tempered basis, random orthonormal MO coefficients" % args.nwater,
        "config": {"workload": "plain FP64 two-step Cholesky of (H2O)%d, O 8s5p3d1f / H 4s2p1d, tau=1e-6, "
                               "+ AO->MO transform nocc=40 nvirt=300" % args.nwater,
                   "nao": basis.nao, "tau": 1e-6, "form": "plain FP64 GPU (the batched, precision and differentiation forms are reported under \"batched\", \"precision\" and \"gradient\")"},
        "cholesky_build_s": t_build,
        "dims": {k: d[k] for k in ("nvecs", "norb_pairs", "nshell_pairs", "nsubsets", "max_subset_dim",
                                   "npivot_shell_pairs", "nmacro")},
        "vectors_gemm_tflops": gemm_flops / tm["vectors_gemm"] / 1e12 if tm["vectors_gemm"] > 0 else None,
        "mo_transform_s": t_mo,
        "nvecs": d["nvecs"],
        "eri_quartets": quartets,
        "phases_s": {k: tm[k] for k in ("vdiag", "step1", "eri_step1", "rtr", "micro", "candidates", "batch",
                                        "step2", "vpqrs",
                                        "potrf_trtri", "vectors", "vabrs", "vectors_gemm", "eri_kernel")},
        "rtr_tflops": tm["rtr_flops"] / tm["rtr"] / 1e12 if tm["rtr"] > 0 else None,
        "roofline": {"bound": "mfma",
                     "kernel": "ERI kernel family: md4_kernel (FP64 4x4x4 MFMA) + md_kernel (FP64 16x16x4 MFMA) + "
                               "md_lane_kernel (FP64 VALU) + k_build_et; MI355X FP64 vector peak = FP64 matrix peak",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None,
                     "traffic": profiled_traffic(),
                     "traffic_source": "profiles/eri_kernels_h2o16_hbm_bytes.md (HBM bytes per launch, "
                                       "FETCH_SIZE + WRITE_SIZE scaled by a copy probe)",
                     "launches": launches, "flops_per_launch": eri_flops / max(launches, 1),
                     "flops": eri_flops, "kernel_s": eri_s},
    }
    stat_rows = sorted(ctx.eri_class_stats(), key=lambda r: -r[3]) if args.eri_stats else []
    # reconstruction error of the headline build (outside the timed region)
    res_max, res_min = ctx.chol_residual()
    head_order = ctx.chol_meta()["pivot_order"]
    out["decode"] = "reference chol2_pq2p_ge_q (TwoStepCholesky_Step2.f90:422-466), reproduced for parity"
    out["residual_max"] = res_max
    out["residual_min"] = res_min
    if not args.no_decode_ties:
        out.update(decode_and_ties(ctx, basis, head_order))
    if not args.no_precision:
        out["precision"] = precision_form(ctx, basis, Cmo, hip)
    ctx.close()
    if not args.no_batched:
        out["batched"] = batched_form(args.batch_workers, not args.no_autotune)
    if not args.no_gradient:
        out["gradient"] = gradient_form()
    if not args.no_thc:
        out["thc"] = thc_form(hip, cpu=not args.no_cpu_baseline)
    if not args.no_contracted:
        out["contracted"] = contracted_form()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.seed)
    if args.eri_stats:
        rows = stat_rows
        tot = sum(r[3] for r in rows)
        print("E F      quartets     ms    TFLOP/s  share", file=sys.stderr)
        for E, F, nq, ms, fl in rows:
            print("%d %d %13d %8.2f %8.3f %6.1f%%" % (E, F, nq, ms, fl / ms / 1e9 if ms else 0,
                                                    100 * ms / tot), file=sys.stderr)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
