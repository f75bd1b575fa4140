#!/usr/bin/env python3
"""Wilson-chain NRG iterations/s on one MI355X.

Workload (the FDM spectral-function configuration): SIAM U=0.3, eps=-0.15, Gamma=0.01 (V=sqrt(2 Gamma/pi)),
Lambda=2, K=2048 kept states, 60 Wilson sites, the observable f^dagger_sigma chi rotated every step,
every truncated iteration recorded for the T=0 FDM backward pass on 2000 log bins.

A step is one Wilson-chain iteration of this chain: add_bath_site (sector assembly + batched
eigensolve), update_system_operator, shell record, update_internal_state (truncation + f-dagger
rotation).  W warm-up iterations run untimed (they include the small pre-truncation iterations),
then exactly K iterations are timed.  Every library call returns with its stream drained, so the
wall clock around the K calls is the device time of the steps.

Beside the timed value the line carries:
  plain_gpu     the same iterations in the plain FP64 GPU form (rocSOLVER dsyevd + rocBLAS dgemm,
                one call per block), a bounded sample; speedup = value / plain
  eig_err       max |lambda_gpu - lambda_LAPACK| over every block of every iteration after the timed ones
  solve_ms_p50/p99  the eigensolve alone (sector eigenpairs of a step), device time per timed step
  spectra       the FDM backward pass over all recorded shells (spectra/s)
  z_average     16 z-shifted chains of the z-averaged thermodynamics workload, one after another and
                stepped together (chain iterations/s)
  rgflow        the RG-flow workload (K=1024, 40 sites) in the mixed and the plain mode
  serving       a Poisson stream of SIAM requests (64 of the 512-request workload), one chain per
                request on 8 host workers: p50/p99 request latency and solves/s
  roofline      the dominant kernel family, from the library's per-launch HIP-event timers
  cpu_baseline  the FP64 oracle (faithful reference form, host LAPACK/BLAS) on host cores
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
CFG = dict(eps=-0.15, U=0.3, Gamma=0.01, Lambda=2.0, K=2048, nsites=60, nbins=2000)


def schedule():
    import nrgplusplus_amd as nk
    L = CFG["Lambda"]
    V = np.sqrt(2 * CFG["Gamma"] / np.pi)
    ts = [V] + [nk.hopping(n, L) for n in range(CFG["nsites"])]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(CFG["nsites"])]
    return ts, rs


class Solve:
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
        if keep_h:
            err = self.eig_error()
            self.g.keep_hamiltonians(False)
        else:
            err = None
        self.g.update_internal_state()
        self.i += 1
        return err

    def eig_error(self):
        qn, dims = self.g.sectors()
        eq = self.g.eigenvaluesQ
        worst = 0.0
        for s, n in enumerate(dims):
            if n:
                w = np.linalg.eigvalsh(self.g.block_h(s))
                worst = max(worst, float(np.abs(w - eq[s]).max()))
        return worst


def cpu_baseline(n_skip, budget_s=20.0):
    import oracle
    cores = len(os.sched_getaffinity(0))
    oracle.set_threads(cores)
    drv = oracle.driver(imp_kind=0, eps=CFG["eps"], U=CFG["U"], bath_kind=0, Lambda=CFG["Lambda"],
                        first_V=np.sqrt(2 * CFG["Gamma"] / np.pi), nsites=CFG["nsites"], K=CFG["K"])
    # one probe iteration sizes the sample to ~budget_s of CPU work
    t1 = oracle.time_forward(drv, n_skip, 1)
    n1 = 2  # then two sites of the same chain at one thread (the reference's serial loop on one core)
    n = int(max(2, min(CFG["nsites"] - n_skip - n1, budget_s / max(t1, 1e-3))))
    secs, secs1 = oracle.time_forward_with_single_core(drv, n_skip, n, n1)
    oracle.set_threads(cores)
    k0 = n_skip + n
    return {"value": n / secs, "unit": "iterations/s", "cores": cores, "kind": "port",
            "sample": "FP64 oracle in the reference's form (per-block LAPACKE_dsyevd + cblas_dgemm via %s, "
                      "update_system_operator, shell save), Wilson sites %d..%d of the same chain"
                      % (oracle.lib().orc_lapack_library().decode(), n_skip, n_skip + n - 1),
            "seconds": secs,
            "single_core": {"value": n1 / secs1, "unit": "iterations/s", "cores": 1,
                            "sample": "Wilson sites %d..%d, one thread" % (k0, k0 + n1 - 1), "seconds": secs1}}


PEAK_F64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0    # MI355X HBM3E


FAMILY_KERNEL = {"jacobi_round": "k_jacobi_round", "gemm_f64": "k_gemm64"}


def measured_traffic(family):
    """HBM bytes per dispatch of the family's kernel from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "pmc_traffic_siam_fdm_K2048.md")
    name = FAMILY_KERNEL.get(family)
    if not name or not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*%s\s*\|.*\|\s*([0-9.eE+-]+)\s*\|\s*$" % name, line)
        if m:
            return float(m.group(1))
    return None


def roofline(kstats):
    """The dominant kernel family of the timed region against its roofline.  The eigensolver
    families are compute kernels (FP64 Gram and FP32 update in the Jacobi rounds, FP64 GEMMs) and are
    priced against the FP64 peak; oa_update, which only streams S, R and E, against HBM."""
    if not kstats:
        return None
    fam, k = max(kstats.items(), key=lambda kv: kv[1]["ms"])
    sec = k["ms"] * 1e-3
    if fam == "oa_update":
        ach = k["bytes"] / sec / 1e9
        return {"kernel": fam, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": None,
                "per_launch": {"bytes": k["bytes"] / k["launches"], "us": 1e3 * k["ms"] / k["launches"]}}
    ach = k["flops"] / sec / 1e12
    return {"kernel": fam, "bound": "mfma", "achieved": ach, "peak": PEAK_F64_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_F64_TFLOPS, "traffic": measured_traffic(fam),
            "algorithmic_bytes_per_launch": k["bytes"] / k["launches"],
            "per_launch": {"flops": k["flops"] / k["launches"], "us": 1e3 * k["ms"] / k["launches"]}}


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


SCFG = dict(requests=64, rate=5.0, workers=8, seed=20261202, nbins=2000, n_omega=1000, b=0.6)


def serving(nk):
    """A bounded sample of the request workload: SIAM requests with U ~ Unif[0.1, 1], eps = -U/2 + d,
    d ~ Unif[-0.1, 0.1], Gamma log-uniform in [0.005, 0.05], Lambda in {2, 2.5, 3}, K in {512, 1024,
    2048}, 40-60 Wilson sites, each returning A(omega) (f-dagger chi, T = 0 FDM, 2000 bins,
    log-Gaussian b = 0.6) on 1000 positive and 1000 negative frequencies.  64 requests (of the
    workload's 512) arrive as a Poisson stream at 5 requests/s (seeded numpy generator); 8 host
    workers each run one request at a time on its own chain and stream.  Latency = completion time -
    arrival time."""
    from concurrent.futures import ThreadPoolExecutor
    c = SCFG
    rng = np.random.default_rng(c["seed"])
    n = c["requests"]
    arrive = np.cumsum(rng.exponential(1.0 / c["rate"], n))
    reqs = []
    for _ in range(n):
        U = rng.uniform(0.1, 1.0)
        reqs.append(dict(U=U, eps=-U / 2 + rng.uniform(-0.1, 0.1), G=float(np.exp(rng.uniform(np.log(0.005), np.log(0.05)))),
                         L=float(rng.choice([2.0, 2.5, 3.0])), K=int(rng.choice([512, 1024, 2048])),
                         nit=int(rng.integers(40, 61))))
    omega = np.logspace(-10, 1, c["n_omega"])

    def solve(r):
        g = nk.NrgCore(nk.Site.spinhalf(r["eps"], r["U"]), nk.Site.spinhalf(0, 0), r["K"])
        op = nk.SystemOperator(g, nk.Site.fdag_chi_spinhalf())
        rec = nk.FdmRecorder(g, op, c["nbins"], 1e-10, 10.0)
        L, V = r["L"], np.sqrt(2 * r["G"] / np.pi)
        g.add_bath_site([V, V], 1.0)
        op.update()
        g.update_internal_state()
        m = 0
        for k in range(r["nit"]):
            t = nk.hopping(k, L)
            g.add_bath_site([t, t], 1.0 if k == 0 else np.sqrt(L))
            op.update()
            if g.checkHigherEnergyDiscarded():
                rec.record()
            else:
                m += 1
            g.update_internal_state()
        pos, neg, _, _ = rec.backward(L, scale_offset=-1, first_cnt=m + 1)
        x = rec.energy_points()
        A = np.concatenate([nk.broaden(x, neg.sum(0), omega, c["b"])[::-1], nk.broaden(x, pos.sum(0), omega, c["b"])])
        return time.perf_counter(), float(pos.sum() + neg.sum()), bool(np.all(np.isfinite(A)))

    solve(dict(U=0.3, eps=-0.15, G=0.01, L=2.0, K=512, nit=12))  # warm-up
    t0 = time.perf_counter()
    futs = []
    with ThreadPoolExecutor(c["workers"]) as ex:
        for k in range(n):
            wait = t0 + arrive[k] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            futs.append(ex.submit(solve, reqs[k]))
        out = [f.result() for f in futs]
    done = np.array([o[0] - t0 for o in out])
    lat = (done - arrive) * 1e3
    return {"workload": "SIAM requests (U, eps, Gamma, Lambda, K, 40-60 sites drawn per request), each returning "
                        "A(omega) on 2000 frequencies; Poisson arrivals",
            "sample": "64 of 512 requests at 5 requests/s, seed %d, 8 workers" % c["seed"],
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "solves_per_s": n / float(done.max() - arrive[0]),
            "total_weight_range": [min(o[1] for o in out), max(o[1] for o in out)], "all_finite": all(o[2] for o in out)}


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


def two_channel(nk, n_steps):
    """The two-channel SIAM chain (impurity spinhalf eps=-0.2, U=0.4 with its f-dagger duplicated to 4
    operators, two-channel bath, Gamma1 = Gamma2 = 0.01, Lambda = 3, K = 4096): the first n_steps
    iterations in the mixed mode (blocks above 768 rows on the rocSOLVER crossover) and in the plain
    FP64 mode; per iteration: states, widest block, ms."""
    L, K, U, eps, G = 3.0, 4096, 0.4, -0.2, 0.01
    V = np.sqrt(2 * G / np.pi)
    out = {}
    for name, mode in (("mixed", nk.NRGK_MODE_MIXED), ("plain_fp64", nk.NRGK_MODE_PLAIN64)):
        g = nk.NrgCore(nk.Site.spinhalf(eps, U).duplicate_ops(), nk.Site.two_channel(), K, mode=mode)
        rows = []
        for n in range(n_steps):
            t = V if n == 0 else nk.hopping(n - 1, L)
            t0 = time.perf_counter()
            g.add_bath_site([t] * 4, 1.0 if n <= 1 else np.sqrt(L))
            g.update_internal_state()
            ms = (time.perf_counter() - t0) * 1e3
            _, dims = g.sectors()
            rows.append({"iteration": n, "states": int(dims.sum()), "widest_block": int(dims.max()), "ms": ms})
        out[name] = rows
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=0, help="0: the timed iterations")
    ap.add_argument("--no-zavg", action="store_true", help="skip the z-averaged multi-chain sample")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-serving sample")
    ap.add_argument("--no-rgflow", action="store_true", help="skip the RG-flow sample")
    ap.add_argument("--two-channel-steps", type=int, default=0,
                    help="also time this many iterations of the two-channel K=4096 chain (mixed and plain)")
    ap.add_argument("--stop-after-timed", action="store_true",
                    help="end after the timed iterations (for counter runs): no eigenvalue check, FDM pass, "
                         "plain or CPU sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU sharding of the NRG step is a design "
              "(independent chains / sector blocks per device) and has no N>1 measurement", file=sys.stderr)
        sys.exit(2)
    total = CFG["nsites"] + 1
    W = max(0, min(args.warmup, total - 1))
    K = max(1, min(args.steps, total - W))

    import nrgplusplus_amd as nk
    nk.lib()
    s = Solve(nk.NRGK_MODE_MIXED)
    for _ in range(W):
        s.step()
    # every library call returns with the chain's stream drained, so the host clock around the K
    # calls is the device time of the steps; the dominant kernel family is timed on that same stream
    # (one event pair per sweep, recorded without blocking)
    s.g.profile(True, only="jacobi_round")
    step_ms = []
    t0 = time.perf_counter()
    solve_ms, unconverged = [], 0
    for _ in range(K):
        ta = time.perf_counter()
        s.step()
        step_ms.append((time.perf_counter() - ta) * 1e3)
        st = s.g.step_stats()  # device time of this step's eigensolve (HIP events around it)
        solve_ms.append(float(st["ms_eig"]))
        unconverged += st["oa_unconverged"]
    t1 = time.perf_counter()
    kstats = s.g.kernel_stats()
    s.g.profile(False)
    ms_per_step = (t1 - t0) * 1e3 / K
    value = K / (t1 - t0)
    if args.stop_after_timed:
        print(json.dumps({"metric": METRIC, "value": value, "unit": "iterations/s", "steps": K, "warmup": W,
                          "ms_per_step": ms_per_step, "kernels": kstats, "partial": "stop-after-timed"}))
        return
    # finish the chain untimed, checking every remaining iteration's eigenvalues against LAPACK
    eig_err, eig_checked = None, 0
    while s.i < total:
        e = s.step(keep_h=True)
        if e is not None:
            eig_err = e if eig_err is None else max(eig_err, e)
            eig_checked += 1
    tb0 = time.perf_counter()
    pos, neg, tr, sums = s.rec.backward(CFG["Lambda"], scale_offset=-1, first_cnt=s.min_it + 1)
    tb1 = time.perf_counter()
    spectra = {"backward_s": tb1 - tb0, "shells": int(len(tr)), "spectra_per_s": 1.0 / (tb1 - tb0),
               "total_weight": float(pos.sum() + neg.sum()), "rho_trace_max_dev": float(np.abs(tr - 1).max())}

    plain = None
    if not args.no_plain:
        p = Solve(nk.NRGK_MODE_PLAIN64, with_fdm=False)
        for _ in range(W):
            p.step()
        n = max(1, min(args.plain_steps or K, total - W))  # default: the same iterations as the timed region
        q0 = time.perf_counter()
        for _ in range(n):
            p.step()
        q1 = time.perf_counter()
        plain = {"value": n / (q1 - q0), "unit": "iterations/s", "steps": n,
                 "form": "rocSOLVER dsyevd + rocBLAS dgemm, one call per block, FP64"}

    zavg = None if args.no_zavg else z_average(nk)
    serve = None if args.no_serving else serving(nk)
    rgf = None if args.no_rgflow else rgflow(nk)
    tc = two_channel(nk, args.two_channel_steps) if args.two_channel_steps > 0 else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W - 1 if W > 0 else 0)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 jacobi + fp64 refinement (results fp64)",
        "data": "synthetic (model Hamiltonian; no dataset)",
        "config": {"workload": "SIAM FDM spectral function: U=0.3 eps=-0.15 Gamma=0.01 Lambda=2 K=2048, "
                               "60 sites, f^dagger chi rotated each step, 2000 log bins",
                   "iterations_timed": [W, W + K - 1]},
        "eig_err": eig_err,
        "eig_err_iterations": [W + K, W + K + eig_checked - 1] if eig_checked else None,
        "plain_gpu": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "spectra": spectra,
        "z_average": zavg,
        "serving": serve,
        "rgflow": rgf,
        "two_channel": tc,
        "step_ms_p50": float(np.percentile(step_ms, 50)),
        "step_ms_p99": float(np.percentile(step_ms, 99)),
        "solve_ms_p50": float(np.percentile(solve_ms, 50)),
        "solve_ms_p99": float(np.percentile(solve_ms, 99)),
        "refinement_unconverged_blocks": unconverged,
        "kernels": {k: dict(ms_per_step=v["ms"] / K, launches_per_step=v["launches"] / K,
                            tflops=(v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 else None,
                            gbs=(v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else None)
                    for k, v in kstats.items()},
        "roofline": roofline(kstats),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
