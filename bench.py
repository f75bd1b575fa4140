#!/usr/bin/env python3
"""Wilson-chain NRG iterations/s on one MI355X.

Headline workload (the largest single-GPU configuration): the two-channel SIAM RG flow.  Impurity
spinhalf eps=-0.2, U=0.4 with its f-dagger duplicated to four operators, two-channel bath (16-state
sites), Gamma1 = Gamma2 = 0.01 (V = sqrt(2 Gamma/pi)), Lambda = 3, K = 4096 kept states, 50 iterations
(first site {V,V,V,V}, then Wilson sites with rescale sqrt(Lambda) from the second), all_eigenvalue
read after every iteration as the RG-flow driver does.  From iteration 3 on every iteration has
~66 000 states in ~100 sector blocks, the widest ~4 200 rows.

A step is one Wilson-chain iteration: add_bath_site (sector assembly + batched eigensolve) and
update_internal_state (truncation + f-dagger rotation).  W warm-up iterations run untimed (the small
pre-truncation iterations), then exactly K iterations are timed.  Every library call returns with its
stream drained, so the wall clock around the K calls is the device time of the steps.

Beside the timed value the line carries:
  plain_gpu     the same iterations in the plain FP64 GPU form (rocSOLVER dsyevd + rocBLAS dgemm,
                one call per block); speedup = value / plain
  eig_err       max |lambda_gpu - lambda_LAPACK| over every block of every timed iteration: the same
                chain replayed untimed with its Hamiltonians kept (bit-identical all_eigenvalue is
                checked), each block against numpy/LAPACK eigvalsh on the same H
  solve_ms_p50/p99  the eigensolve alone (sector eigenpairs of a step), device time per timed step
  eigensolve_roofline  F_eig = sum_b (4/3) n_b^3 + 2 n_b^2 k_b of the timed iterations over the
                eigensolve's device time, against the FP32 matrix peak
  roofline      the dominant kernel, k_ht_symv (the symmetric products of the wide blocks'
                tridiagonalisation), from HIP events on its stream around one launch in eight over the
                timed steps: algorithmic bytes (the trailing lower triangles read once) per launch over
                the average launch duration, against HBM; traffic from the committed rocprof counters
  kernels       per kernel family device time, launches and rates, from a replay of the timed
                iterations
  siam_fdm      the SIAM FDM spectral-function workload (K=2048, 60 sites, 2000 bins): mixed vs plain,
                eigenvalue error, backward pass, HBM fractions of assembly and binning
  z_average     16 z-shifted chains of the z-averaged thermodynamics workload, one after another and
                stepped together (chain iterations/s)
  rgflow        the SIAM RG-flow workload (K=1024, 40 sites) in the mixed and the plain mode
  serving       a Poisson stream of SIAM requests: p50/p99 request latency and solves/s
  cpu_baseline  the FP64 oracle (the reference's form, host LAPACK/BLAS) on host cores
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NRG iterations/s at <=1e-10 eig err, batched-mixed vs plain GPU; p50/p99 solve ms"
# headline: the two-channel SIAM RG flow
TCFG = dict(eps=-0.2, U=0.4, Gamma=0.01, Lambda=3.0, K=4096, iterations=50)
# secondary: the SIAM FDM spectral function
CFG = dict(eps=-0.15, U=0.3, Gamma=0.01, Lambda=2.0, K=2048, nsites=60, nbins=2000)

PEAK_F32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_F64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0    # MI355X HBM3E


class TwoChannel:
    """The two-channel RG-flow chain stepped one iteration at a time."""

    def __init__(self, nk, mode):
        c = TCFG
        self.g = nk.NrgCore(nk.Site.spinhalf(c["eps"], c["U"]).duplicate_ops(), nk.Site.two_channel(), c["K"], mode=mode)
        L = c["Lambda"]
        V = np.sqrt(2 * c["Gamma"] / np.pi)
        self.ts = [V] + [nk.hopping(n, L) for n in range(c["iterations"] - 1)]
        self.rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(c["iterations"] - 1)]
        self.i = 0

    def step(self):
        self.g.add_bath_site([self.ts[self.i]] * 4, self.rs[self.i])
        self.g.update_internal_state()
        self.i += 1
        return self.g.all_eigenvalue


def f_eig(dims, kept):
    """SURVEY's nominal eigensolve work of one iteration (RG flow: eigenvectors of the kept states)"""
    n = np.asarray(dims, float)
    k = np.asarray(kept, float)
    return float(np.sum(4.0 / 3.0 * n ** 3 + 2.0 * n ** 2 * k))


def eig_error_of_step(g):
    """max over the blocks of the step just assembled of |lambda_gpu - lambda_LAPACK| on the same H"""
    _, dims = g.sectors()
    eq = g.eigenvaluesQ
    worst = 0.0
    for s, n in enumerate(dims):
        if n:
            w = np.linalg.eigvalsh(g.block_h(s))
            worst = max(worst, float(np.abs(w - eq[s]).max()))
    return worst


def progress(msg):
    """a progress line on stderr (the JSON line alone goes to stdout)"""
    print("bench: " + msg, file=sys.stderr, flush=True)


def kernel_table(kstats, K):
    return {k: dict(ms_per_step=v["ms"] / K, launches_per_step=v["launches"] / K,
                    us_per_launch=1e3 * v["ms"] / max(1, v["launches"]),
                    tflops=(v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 and v["flops"] > 0 else None,
                    gbs=(v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 and v["bytes"] > 0 else None)
            for k, v in kstats.items()}


# compute families (flops counted) and HBM families (bytes counted) of the library's timers
HBM_FAMILIES = ("oa_update", "assemble", "fdm_bin", "ht_symv", "tridiagonalise", "tridiag_vectors")
F32_FAMILIES = ("jacobi_round",)


def measured_traffic(kernel):
    """HBM bytes per dispatch of a kernel from the committed counter summaries under profiles/
    (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, calibrated on a copy of known size), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "pmc_traffic_*.md"))):
        for line in open(path):
            m = re.match(r"\|\s*%s\s*\|.*\|\s*([0-9.eE+-]+)\s*\|\s*$" % kernel, line)
            if m:
                return float(m.group(1)), os.path.basename(path)
    return None, None


FAMILY_KERNEL = {"jacobi_round": "k_jacobi_round", "gemm_f64": "k_gemm64", "ht_symv": "k_ht_symv",
                 "assemble": "k_assemble", "fdm_bin": "k_fdm_bin"}


def roofline_of(fam, k):
    """one kernel family against its roof: algorithmic bytes (HBM families) or flops per launch over
    the average launch duration"""
    sec = k["ms"] * 1e-3
    us = 1e3 * k["ms"] / max(1, k["launches"])
    traffic, src = measured_traffic(FAMILY_KERNEL.get(fam, ""))
    if fam in HBM_FAMILIES or k["flops"] <= 0:
        ach = k["bytes"] / sec / 1e9
        return {"kernel": fam, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": traffic, "traffic_source": src,
                "per_launch": {"bytes": k["bytes"] / max(1, k["launches"]), "us": us}}
    peak = PEAK_F32_TFLOPS if fam in F32_FAMILIES else PEAK_F64_TFLOPS
    ach = k["flops"] / sec / 1e12
    return {"kernel": fam, "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
            "frac": ach / peak, "traffic": traffic, "traffic_source": src,
            "algorithmic_bytes_per_launch": k["bytes"] / max(1, k["launches"]),
            "per_launch": {"flops": k["flops"] / max(1, k["launches"]), "us": us}}


def roofline(kstats):
    if not kstats:
        return None
    fam, k = max(kstats.items(), key=lambda kv: kv[1]["ms"])
    return roofline_of(fam, k)


def two_channel_headline(nk, W, K, with_plain=True, with_check=True):
    """The headline: W untimed iterations, K timed, then the plain FP64 form over the same iterations
    and the eigenvalue check on a replay of the timed iterations."""
    s = TwoChannel(nk, nk.NRGK_MODE_MIXED)
    for i in range(W):
        s.step()
        progress("two-channel warm-up iteration %d done" % i)
    # the dominant kernel, k_ht_symv, timed by HIP events on its stream around one launch in eight
    s.g.profile(True, only="ht_symv")
    step_ms, solve_ms, feig, unconv, launches, syncs, eigs, ties = [], [], 0.0, 0, [], [], [], []
    t0 = time.perf_counter()
    for _ in range(K):
        ta = time.perf_counter()
        e = s.step()
        step_ms.append((time.perf_counter() - ta) * 1e3)
        progress("two-channel timed iteration %d: %.1f ms" % (W + len(step_ms) - 1, step_ms[-1]))
        eigs.append(e)
        st = s.g.step_stats()
        solve_ms.append(float(st["ms_eig"]))
        unconv += st["oa_unconverged"]
        launches.append(st["launches"])
        syncs.append(st["host_syncs"])
        ties.append(st["near_ties"])
    t1 = time.perf_counter()
    kstats = s.g.kernel_stats()
    s.g.profile(False)
    value = K / (t1 - t0)
    progress("two-channel mixed: %d timed iterations, %.1f ms/step" % (K, (t1 - t0) * 1e3 / K))
    out = {"value": value, "ms_per_step": (t1 - t0) * 1e3 / K, "step_ms": step_ms, "solve_ms": solve_ms,
           "refinement_unconverged_blocks": unconv, "launches_per_step": float(np.mean(launches)),
           "host_syncs_per_step": float(np.mean(syncs)), "near_ties": int(np.sum(ties)), "kstats": kstats}
    # F_eig of the timed iterations (block dims and kept counts, from a replay with the Hamiltonians kept)
    if with_check:
        r = TwoChannel(nk, nk.NRGK_MODE_MIXED)
        for _ in range(W):
            r.step()
        r.g.profile(True)  # every kernel family, on the replay
        err, same, widest = 0.0, True, 0
        for k in range(K):
            r.g.keep_hamiltonians(True)
            r.g.add_bath_site([r.ts[r.i]] * 4, r.rs[r.i])
            err = max(err, eig_error_of_step(r.g))
            _, dims = r.g.sectors()
            widest = max(widest, int(np.max(dims)))
            r.g.keep_hamiltonians(False)
            r.g.update_internal_state()
            r.i += 1
            feig += f_eig(dims, r.g.kept_counts())
            a = r.g.all_eigenvalue
            same = same and len(a) == len(eigs[k]) and bool(np.array_equal(a, eigs[k]))
        out.update(eig_err=err, replay_bit_identical=same, widest_block=widest, F_eig=feig,
                   kstats_replay=r.g.kernel_stats())
        r.g.profile(False)
        progress("two-channel check: eig_err %.3g, replay identical %s" % (err, same))
    if with_plain:
        p = TwoChannel(nk, nk.NRGK_MODE_PLAIN64)
        for _ in range(W):
            p.step()
        q0 = time.perf_counter()
        pe = None
        for _ in range(K):
            pe = p.step()
        q1 = time.perf_counter()
        progress("two-channel plain: %.1f ms/step" % ((q1 - q0) * 1e3 / K))
        out["plain"] = {"value": K / (q1 - q0), "unit": "iterations/s", "steps": K,
                        "form": "rocSOLVER dsyevd + rocBLAS dgemm, one call per block, FP64",
                        "final_all_eigenvalue_max_diff_vs_mixed":
                            float(np.abs(pe - eigs[-1]).max()) if len(pe) == len(eigs[-1]) else None}
    return out


def cpu_baseline_two_channel(W):
    """The oracle on the host cores in the RG-flow driver form: iteration W (the first timed one, a
    full ~66 000-state iteration) after the untimed iterations before it."""
    import oracle
    cores = len(os.sched_getaffinity(0))
    oracle.set_threads(cores)
    c = TCFG
    drv = oracle.driver(imp_kind=0, eps=c["eps"], U=c["U"], bath_kind=1, dup_ops=1, Lambda=c["Lambda"],
                        first_V=float(np.sqrt(2 * c["Gamma"] / np.pi)), nsites=W, K=c["K"])
    secs = oracle.time_rgflow(drv, W - 1, 1)  # loop site W-1 is iteration W
    return {"value": 1.0 / secs, "unit": "iterations/s", "cores": cores, "kind": "port",
            "sample": "FP64 oracle in the reference's form (per-block LAPACKE_dsyevd + cblas_dgemm via %s), "
                      "two-channel K=4096 iteration %d (the first timed one) after iterations 0..%d untimed"
                      % (oracle.lib().orc_lapack_library().decode(), W, W - 1),
            "seconds": secs}


def cpu_baseline(n_skip, budget_s=20.0):
    """the oracle on the SIAM FDM workload (secondary line)"""
    import oracle
    cores = len(os.sched_getaffinity(0))
    oracle.set_threads(cores)
    drv = oracle.driver(imp_kind=0, eps=CFG["eps"], U=CFG["U"], bath_kind=0, Lambda=CFG["Lambda"],
                        first_V=np.sqrt(2 * CFG["Gamma"] / np.pi), nsites=CFG["nsites"], K=CFG["K"])
    t1 = oracle.time_forward(drv, n_skip, 1)
    n = int(max(2, min(CFG["nsites"] - n_skip, budget_s / max(t1, 1e-3))))
    secs = oracle.time_forward(drv, n_skip, n)
    return {"value": n / secs, "unit": "iterations/s", "cores": cores, "kind": "port",
            "sample": "FP64 oracle (per-block LAPACKE_dsyevd + cblas_dgemm via %s, update_system_operator, shell "
                      "save), SIAM FDM Wilson sites %d..%d" % (oracle.lib().orc_lapack_library().decode(), n_skip,
                                                              n_skip + n - 1), "seconds": secs}


def schedule():
    import nrgplusplus_amd as nk
    L = CFG["Lambda"]
    V = np.sqrt(2 * CFG["Gamma"] / np.pi)
    ts = [V] + [nk.hopping(n, L) for n in range(CFG["nsites"])]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(CFG["nsites"])]
    return ts, rs


class Solve:
    """the SIAM FDM chain: add_bath_site, update_system_operator, shell record, update_internal_state"""

    def __init__(self, mode, with_fdm=True):
        import nrgplusplus_amd as nk
        self.nk = nk
        self.g = nk.NrgCore(nk.Site.spinhalf(CFG["eps"], CFG["U"]), nk.Site.spinhalf(0, 0), CFG["K"], mode=mode)
        self.op = nk.SystemOperator(self.g, nk.Site.fdag_chi_spinhalf())
        self.rec = nk.FdmRecorder(self.g, self.op, CFG["nbins"], 1e-10, 10.0) if with_fdm else None
        self.ts, self.rs = schedule()
        self.i = 0
        self.min_it = 0

    def step(self, keep_h=False):
        t, r = self.ts[self.i], self.rs[self.i]
        if keep_h:
            self.g.keep_hamiltonians(True)
        self.g.add_bath_site([t, t], r)
        self.op.update()
        if self.i > 0:
            if self.g.checkHigherEnergyDiscarded():
                if self.rec is not None:
                    self.rec.record()
            else:
                self.min_it += 1
        err = eig_error_of_step(self.g) if keep_h else None
        if keep_h:
            self.g.keep_hamiltonians(False)
        self.g.update_internal_state()
        self.i += 1
        return err


def siam_fdm(nk, W=10, K=30, with_plain=True, with_cpu=True):
    """The SIAM FDM workload: W untimed, K timed iterations (mixed), the rest of the chain with every
    block checked against LAPACK, the backward pass, the plain FP64 form over the timed iterations."""
    total = CFG["nsites"] + 1
    s = Solve(nk.NRGK_MODE_MIXED)
    for _ in range(W):
        s.step()
    s.g.profile(True)
    t0 = time.perf_counter()
    for _ in range(K):
        s.step()
    t1 = time.perf_counter()
    kst = s.g.kernel_stats()
    s.g.profile(False)
    value = K / (t1 - t0)
    eig_err = 0.0
    while s.i < total:
        eig_err = max(eig_err, s.step(keep_h=True))
    s.g.profile(True, only="fdm_bin")
    tb0 = time.perf_counter()
    pos, neg, tr, sums = s.rec.backward(CFG["Lambda"], scale_offset=-1, first_cnt=s.min_it + 1)
    tb1 = time.perf_counter()
    bst = s.g.kernel_stats()
    s.g.profile(False)
    out = {"workload": "SIAM FDM spectral function: U=0.3 eps=-0.15 Gamma=0.01 Lambda=2 K=2048, 60 sites, "
                       "f^dagger chi rotated each step, 2000 log bins",
           "iterations_timed": [W, W + K - 1], "value": value, "unit": "iterations/s", "ms_per_step": (t1 - t0) * 1e3 / K,
           "eig_err_untimed_iterations": eig_err, "eig_err_iterations": [W + K, total - 1],
           "spectra": {"backward_s": tb1 - tb0, "shells": int(len(tr)), "spectra_per_s": 1.0 / (tb1 - tb0),
                       "total_weight": float(pos.sum() + neg.sum()), "signed_weights": s.rec.signed_weights().tolist(),
                       "rho_trace_max_dev": float(np.abs(tr - 1).max())},
           "kernels": kernel_table(kst, K),
           "roofline": roofline(kst)}
    for fam, st in (("assemble", kst.get("assemble")), ("fdm_bin", bst.get("fdm_bin"))):
        if st:
            out["hbm_" + fam] = roofline_of(fam, st)
    if with_plain:
        p = Solve(nk.NRGK_MODE_PLAIN64, with_fdm=False)
        for _ in range(W):
            p.step()
        q0 = time.perf_counter()
        for _ in range(K):
            p.step()
        q1 = time.perf_counter()
        out["plain_gpu"] = {"value": K / (q1 - q0), "unit": "iterations/s", "steps": K}
        out["speedup_vs_plain"] = value / out["plain_gpu"]["value"]
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline(W - 1 if W > 0 else 0, budget_s=10.0)
    return out


ZCFG = dict(eps=-0.25, U=0.5, Gamma=0.015, Lambda=4.0, K=1024, nsites=45, nz=16, nT=200)


def z_average(nk):
    """The z-averaged thermodynamics workload: SIAM U=0.5, eps=-0.25, Gamma=0.015, Lambda=4, K=1024,
    45 sites, one chain per twist z = k/16 (z = 0 taken as z = 1, the same mesh) with the hoppings of
    the z-shifted discretisation, entropy and specific heat on 200 log-spaced T in [1e-9, 1], averaged
    over z.  Timed twice: the 16 chains one after another, and stepped together (nrgk_chains_step:
    one host worker and stream per chain).  Both must give the same entropy to the last bit."""
    c = ZCFG
    L, ns, nz = c["Lambda"], c["nsites"], c["nz"]
    A = 0.5 * np.log(L) * (1 + 1 / L) / (1 - 1 / L)
    zs = [(k if k else nz) / nz for k in range(nz)]
    hops = [nk.z_chain(z, L, ns)[0] * A for z in zs]
    V = np.sqrt(2 * c["Gamma"] / np.pi)
    T = np.logspace(-9, 0, c["nT"])

    def new_chain():
        g = nk.NrgCore(nk.Site.spinhalf(c["eps"], c["U"]), nk.Site.spinhalf(0, 0), c["K"])
        return g, nk.FdmRecorder(g, None)

    def alone(k):
        g, rec = new_chain()
        g.add_bath_site([V, V], 1.0)
        g.update_internal_state()
        m = 0
        for n in range(ns):
            g.add_bath_site([hops[k][n]] * 2, 1.0 if n == 0 else np.sqrt(L))
            if g.checkHigherEnergyDiscarded():
                rec.record()
            else:
                m += 1
            g.update_internal_state()
        return rec.thermodynamics(L, T, scale_offset=-1, first_cnt=m + 1)[0]

    alone(0)  # warm-up (first allocations, kernel loads)
    t0 = time.perf_counter()
    S_seq = np.mean([alone(k) for k in range(nz)], axis=0)
    t1 = time.perf_counter()
    gs, recs = zip(*[new_chain() for _ in range(nz)])
    nk.chains_step(gs, [[V, V]] * nz, 1.0)
    m = np.zeros(nz, int)
    for n in range(ns):
        m += ~nk.chains_step(gs, [[h[n]] * 2 for h in hops], 1.0 if n == 0 else np.sqrt(L), recorders=recs)
    S_tog = np.mean([recs[k].thermodynamics(L, T, scale_offset=-1, first_cnt=int(m[k]) + 1)[0]
                     for k in range(nz)], axis=0)
    t2 = time.perf_counter()
    its = nz * (ns + 1)
    return {"workload": "z-averaged SIAM thermodynamics: U=0.5 eps=-0.25 Gamma=0.015 Lambda=4 K=1024, 45 sites, "
                        "16 twists z=k/16, 200 temperatures",
            "chains_together_s": t2 - t1, "chains_one_after_another_s": t1 - t0,
            "chain_iterations_per_s": its / (t2 - t1), "speedup_together": (t1 - t0) / (t2 - t1),
            "identical": bool(np.array_equal(S_seq, S_tog))}


SCFG = dict(requests=512, seed=20261202, nbins=2000, n_omega=1000, b=0.6, slots=8, probe=48, load=0.8)


def serving(nk):
    """The request workload through the library's request loop (nrgk_server_*): SIAM requests with
    U ~ Unif[0.1, 1], eps = -U/2 + d, d ~ Unif[-0.1, 0.1], Gamma log-uniform in [0.005, 0.05], Lambda in
    {2, 2.5, 3}, K in {512, 1024, 2048}, 40-60 Wilson sites, each returning A(omega) (f-dagger chi, T = 0
    FDM, 2000 bins, log-Gaussian b = 0.6) at 1000 positive and 1000 negative frequencies.  8 slots,
    admission by device footprint.  First the capacity: a burst of 48 requests submitted at once
    (solves/s while saturated).  Then all 512 requests as a Poisson stream (seeded numpy generator) at
    0.8 of that capacity; latency = completion - arrival."""
    c = SCFG
    rng = np.random.default_rng(c["seed"])
    n = c["requests"]
    reqs = []
    for _ in range(n):
        U = rng.uniform(0.1, 1.0)
        reqs.append(dict(U=U, eps=-U / 2 + rng.uniform(-0.1, 0.1), G=float(np.exp(rng.uniform(np.log(0.005), np.log(0.05)))),
                         L=float(rng.choice([2.0, 2.5, 3.0])), K=int(rng.choice([512, 1024, 2048])),
                         nit=int(rng.integers(40, 61))))
    omega = np.logspace(-10, 1, c["n_omega"])

    def submit(srv, r):
        return srv.submit(nk.Site.spinhalf(r["eps"], r["U"]), nk.Site.spinhalf(0, 0), nk.Site.fdag_chi_spinhalf(),
                          r["K"], r["L"], float(np.sqrt(2 * r["G"] / np.pi)), r["nit"], c["nbins"], omega, c["b"])

    srv = nk.Server(n_slots=c["slots"])
    srv.wait(submit(srv, dict(U=0.3, eps=-0.15, G=0.01, L=2.0, K=512, nit=12)))  # warm-up
    # capacity: a saturated burst
    t0 = time.perf_counter()
    ids = [submit(srv, reqs[k % n]) for k in range(c["probe"])]
    for i in ids:
        srv.wait(i)
    cap = c["probe"] / (time.perf_counter() - t0)
    progress("serving capacity %.2f solves/s" % cap)
    rate = c["load"] * cap
    arrive = np.cumsum(np.random.default_rng(c["seed"] + 1).exponential(1.0 / rate, n))
    t0 = time.perf_counter()
    ids, sub_t = [], []
    for k in range(n):
        wait = t0 + arrive[k] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        sub_t.append(time.perf_counter() - t0)
        ids.append(submit(srv, reqs[k]))
        if k % 64 == 63:
            progress("serving: %d of %d requests submitted" % (k + 1, n))
    lat, qms, tw, fin = [], [], [], []
    for i in ids:
        ap, an, info = srv.wait(i)
        lat.append(info["latency_ms"])
        qms.append(info["queue_ms"])
        tw.append(info["total_weight"])
        fin.append(bool(np.all(np.isfinite(ap)) and np.all(np.isfinite(an))))
    t1 = time.perf_counter()
    st = srv.stats()
    srv.close()
    # latency from the scheduled arrival: submission lag of the host loop included
    lat = np.array(lat) + 1e3 * (np.array(sub_t) - arrive)
    return {"workload": "SIAM requests (U, eps, Gamma, Lambda, K, 40-60 sites drawn per request), each returning "
                        "A(omega) on 2000 frequencies; Poisson arrivals",
            "server": "nrgk_server: %d slots, admission by device footprint, one chain and stream per request" % c["slots"],
            "capacity_solves_per_s": cap,
            "sample": "all %d requests at %.2f requests/s (%.1f x capacity), seed %d" % (n, rate, c["load"], c["seed"]),
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "queue_ms_p50": float(np.percentile(qms, 50)), "queue_ms_p99": float(np.percentile(qms, 99)),
            "solves_per_s": n / (t1 - t0), "peak_running": st["peak_running"],
            "total_weight_range": [float(min(tw)), float(max(tw))], "all_finite": all(fin)}


def rgflow(nk):
    """The RG-flow workload (examples/rgflowSIAM form): spinhalf impurity eps=-0.25, U=0.5, Gamma=0.02,
    Lambda=2.5, K=1024, first site {V, V} then 40 Wilson sites; all_eigenvalue read after every
    iteration.  The whole flow in the mixed and in the plain FP64 mode (each after one untimed flow),
    and the largest difference of the final all_eigenvalue between them."""
    L, K, ns = 2.5, 1024, 40
    V = np.sqrt(2 * 0.02 / np.pi)

    def flow(mode):
        g = nk.NrgCore(nk.Site.spinhalf(-0.25, 0.5), nk.Site.spinhalf(0, 0), K, mode=mode)
        t0 = time.perf_counter()
        g.add_bath_site([V, V], 1.0)
        g.update_internal_state()
        e = g.all_eigenvalue
        for n in range(ns):
            t = nk.hopping(n, L)
            g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(L))
            g.update_internal_state()
            e = g.all_eigenvalue
        return time.perf_counter() - t0, e

    out = {"workload": "SIAM RG flow: U=0.5 eps=-0.25 Gamma=0.02 Lambda=2.5 K=1024, 40 sites"}
    res = {}
    for name, mode in (("mixed", nk.NRGK_MODE_MIXED), ("plain_fp64", nk.NRGK_MODE_PLAIN64)):
        flow(mode)
        secs, e = flow(mode)
        res[name] = e
        out[name] = {"seconds": secs, "iterations_per_s": (ns + 1) / secs}
    out["speedup"] = out["plain_fp64"]["seconds"] / out["mixed"]["seconds"]
    a, b = res["mixed"], res["plain_fp64"]
    out["final_all_eigenvalue_max_diff"] = float(np.abs(a - b).max()) if len(a) == len(b) else None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the eigenvalue check (replay of the timed steps)")
    ap.add_argument("--no-fdm", action="store_true", help="skip the SIAM FDM workload line")
    ap.add_argument("--no-zavg", action="store_true", help="skip the z-averaged multi-chain sample")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-serving sample")
    ap.add_argument("--no-rgflow", action="store_true", help="skip the RG-flow sample")
    ap.add_argument("--stop-after-timed", action="store_true",
                    help="end after the timed iterations (for profiler runs): no check, plain, samples or CPU leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU sharding of the NRG step is a design "
              "(independent chains / sector blocks per device) and has no N>1 measurement", file=sys.stderr)
        sys.exit(2)
    total = TCFG["iterations"]
    W = max(1, min(args.warmup, total - 1))
    K = max(1, min(args.steps, total - W))

    import nrgplusplus_amd as nk
    nk.lib()
    quick = args.stop_after_timed
    h = two_channel_headline(nk, W, K, with_plain=not (args.no_plain or quick), with_check=not (args.no_check or quick))
    kst = h.pop("kstats")
    kst_all = h.pop("kstats_replay", None)
    line = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 (eigensolve: fp32 jacobi + fp64 refinement below the crossover, fp64 householder above)",
        "data": "synthetic (model Hamiltonian; no dataset)",
        "config": {"workload": "two-channel SIAM RG flow: U=0.4 eps=-0.2 Gamma1=Gamma2=0.01 Lambda=3 K=4096, "
                               "50 iterations, all_eigenvalue read each iteration",
                   "iterations_timed": [W, W + K - 1], "widest_block": h.get("widest_block")},
        "eig_err": h.get("eig_err"),
        "eig_err_iterations": [W, W + K - 1] if "eig_err" in h else None,
        "replay_bit_identical": h.get("replay_bit_identical"),
        "plain_gpu": h.get("plain"),
        "speedup_vs_plain": (h["value"] / h["plain"]["value"]) if h.get("plain") else None,
        "step_ms_p50": float(np.percentile(h["step_ms"], 50)),
        "step_ms_p99": float(np.percentile(h["step_ms"], 99)),
        "solve_ms_p50": float(np.percentile(h["solve_ms"], 50)),
        "solve_ms_p99": float(np.percentile(h["solve_ms"], 99)),
        "launches_per_step": h["launches_per_step"],
        "host_syncs_per_step": h["host_syncs_per_step"],
        "refinement_unconverged_blocks": h["refinement_unconverged_blocks"],
        "truncation_near_ties": h["near_ties"],
        "kernels": kernel_table(kst_all, K) if kst_all else None,
        "roofline": roofline_of("ht_symv", kst["ht_symv"]) if "ht_symv" in kst else None,
    }
    if line["roofline"] is not None:
        line["roofline"]["sampled"] = "k_ht_symv launches of one column in 8, timed region"
    if "F_eig" in h:
        sec = sum(h["solve_ms"]) * 1e-3
        line["eigensolve_roofline"] = {"F_eig_per_step": h["F_eig"] / K, "eigensolve_s_per_step": sec / K,
                                       "achieved": h["F_eig"] / sec / 1e12, "peak": PEAK_F32_TFLOPS,
                                       "unit": "TFLOP/s", "frac": h["F_eig"] / sec / 1e12 / PEAK_F32_TFLOPS}
    if not quick:
        if not args.no_fdm:
            line["siam_fdm"] = siam_fdm(nk, with_plain=not args.no_plain, with_cpu=not args.no_cpu)
            progress("siam fdm done")
        line["z_average"] = None if args.no_zavg else z_average(nk)
        progress("z-average done")
        line["serving"] = None if args.no_serving else serving(nk)
        progress("serving done")
        line["rgflow"] = None if args.no_rgflow else rgflow(nk)
        progress("rgflow done")
        line["cpu_baseline"] = None if args.no_cpu else cpu_baseline_two_channel(W)
        progress("cpu baseline done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
