"""ESS/s of one VMC step (sampling + E_loc + O_k + SR solve) on one MI355X.

Workload (default): 2D J1 Heisenberg 10x10 periodic (N=100, 600 Pauli strings),
complex RBM with biases alpha=4 (PsiDeep, M=400, P=40,500, seeded complex
Gaussian sigma=0.01, final weights 1), 16,384 persistent chains, 100 exchange
sweeps per step in the S_z=0 sector, angles on MFMA (bf16x3 split, fp32
accumulation), fp32 sweeps / E_loc / O_k, fp64 reductions and a matrix-free CG
solve of (S + 1e-3 I) x = F with fp64 vectors, then the SR parameter update.

ESS per step = Ns * min(1, S_sw / (2 tau_int)), tau_int (in sweeps) of E_loc
measured on 1024 chains recording E_loc every sweep (Sokal window c = 5).
Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def tau_int_sokal(series, c=5.0):
    """integrated autocorrelation time (Sokal convention, tau=1/2 for iid) of
    series[T, chains], autocorrelation averaged over chains"""
    x = series - series.mean(axis=0, keepdims=True)
    T = x.shape[0]
    f = np.fft.rfft(x, n=2 * T, axis=0)
    acf = np.fft.irfft(f * np.conj(f), axis=0)[:T].mean(axis=1)
    if acf[0] <= 0:
        return 0.5
    acf = acf / acf[0]
    tau = 0.5
    for W in range(1, T):
        tau += acf[W]
        if W >= c * tau:
            break
    return max(tau, 0.5)


def tau_batch_means(series, batches=20):
    """tau from batch means: var(batch mean) * b / (2 var)"""
    T = series.shape[0] // batches * batches
    x = series[:T]
    b = T // batches
    bm = x.reshape(batches, b, -1).mean(axis=1)
    v = x.var()
    if v <= 0:
        return 0.5
    return max(0.5, b * bm.var(axis=0, ddof=1).mean() / (2 * v))


def build_workload(A, L=10, alpha=4, seed=3, precision="fp32"):
    N = L * L
    psi = A.seeded_rbm_deep(N, alpha * N, seed=seed, sigma=0.01, final_weight=1.0)
    psi.precision = precision
    H = A.hamiltonians.heisenberg_square(L)
    return N, psi, H


def cpu_baseline(A, L, alpha, seed, seconds):
    """Oracle (fp64 C restatement of the reference host path, OpenMP over chains)
    on a bounded sample of the same workload: chains x 100 exchange sweeps, then
    E_loc and O_k per sample.  Excludes the SR solve (dense S of 40,500^2 does not
    fit a CPU step), which makes the baseline optimistic."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    N, psi, H = build_workload(A, L, alpha, seed, "fp64")
    o = orc.deep(N, psi.a, psi.b, psi.connections, psi.W, psi.final_weights, 0.0)
    terms = list(H)
    words = 2
    a = np.zeros((len(terms), words), dtype=np.uint64)
    b = np.zeros((len(terms), words), dtype=np.uint64)
    cc = np.zeros(len(terms), dtype=np.complex128)
    for k, ((ma, mb), coef) in enumerate(terms):
        for w in range(words):
            a[k, w] = (ma >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
            b[k, w] = (mb >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
        cc[k] = coef
    op = orc.OracleOperator(a, b, cc, words)
    cores = os.cpu_count() or 1
    # calibrate on 4 chains per core x 10 sweeps, then size the sample to ~`seconds`
    t0 = time.perf_counter()
    orc.mc_run(o, 4 * cores, 1, 10, 0, move=orc.MOVE_EXCHANGE, seed=1, num_threads=cores)
    per_sweep = (time.perf_counter() - t0) / 40          # `cores` chains in parallel
    chains = int(seconds / (100 * per_sweep * 1.2) * cores)
    chains = max(cores, min(4096, chains // cores * cores))
    t0 = time.perf_counter()
    r = orc.mc_run(o, chains, 1, 100, 0, move=orc.MOVE_EXCHANGE, seed=2, num_threads=cores)
    w = np.full(chains, 1.0 / chains)
    orc.tdvp_accumulate(op, o, r["samples"], w, with_S=False, num_threads=cores)
    dt = time.perf_counter() - t0
    return dict(samples_per_s=chains / dt, seconds=dt, chains=chains, cores=cores)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=10)
    ap.add_argument("--alpha", type=int, default=4)
    ap.add_argument("--chains", type=int, default=16384)
    ap.add_argument("--sweeps", type=int, default=100)
    ap.add_argument("--cg-iters", type=int, default=100)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import annongpu_amd as A
    A.setDevice(0)
    N, psi, H = build_workload(A, args.L, args.alpha)
    op = A.Operator(H)
    B = args.chains
    ens = A.MonteCarloSpins(B, args.sweeps, 20, B, move="exchange", seed=17, persistent=True)
    tdvp = A.TDVP(psi.num_params)
    shift = 1e-3

    def step():
        tdvp.eval_F(op, psi, ens)
        x = tdvp.solve(shift, "cg", rtol=1e-4, max_iter=args.cg_iters, raise_on_failure=False)
        psi.params = psi.params - args.lr * x

    for _ in range(args.warmup):
        step()
    A.profiler_reset()
    A.start_profiling()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    A.stop_profiling()     # synchronises the device
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    sweep_ms, sweep_n = A.profiler_query("mc_sweep")
    kernel_ms = {k: A.profiler_query(k) for k in ("mc_sweep", "theta", "local_energy", "o_k", "spin_sums",
                                                  "sr_matvec")}

    # autocorrelation of E_loc, one sweep apart, on 1024 chains of the same psi
    Bt, T = 1024, 400
    ens_t = A.MonteCarloSpins(Bt * T, 1, 50, Bt, move="exchange", seed=23)
    el = A.local_energies(op, psi, ens_t).real.reshape(T, Bt)
    tau = tau_int_sokal(el)
    tau_bm = tau_batch_means(el)
    ess_frac = min(1.0, args.sweeps / (2.0 * tau))
    ess_step = B * ess_frac
    value = ess_step / (ms_step * 1e-3)

    # roofline of the dominant kernel: exchange sweeps, 36 flop per hidden unit per proposal
    M = args.alpha * N
    flops_launch = B * args.sweeps * N * 36.0 * M
    avg_ms = sweep_ms / max(sweep_n, 1)
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12
    peak = 157.3

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(A, args.L, args.alpha, 3, args.cpu_seconds)
        cpu = {"value": c["samples_per_s"] * ess_frac, "unit": "ESS/s", "cores": c["cores"], "kind": "port",
               "sample": f"{c['chains']} chains x {args.sweeps} exchange sweeps + E_loc + O_k in fp64 "
                         f"({c['seconds']:.1f} s, OpenMP over chains); SR solve excluded; ESS fraction "
                         f"of the GPU run applied"}

    line = {
        "metric": "ESS/s with E_loc+O_k+SR, RBM N=100 alpha=4: learned-proposal bf16 vs plain fp32",
        "value": value, "unit": "ESS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"Heisenberg {args.L}x{args.L} PBC, complex RBM alpha={args.alpha} "
                               f"(P={psi.num_params}), {B} chains x {args.sweeps} exchange sweeps, "
                               "plain sampler, fp32 + bf16x3 angle GEMM, CG SR (fp64 dots)",
                   "sampler": "plain-exchange", "tau_int_sweeps": tau, "tau_batch_means": tau_bm,
                   "ess_per_step": ess_step, "samples_per_step": B, "acceptance": ens.acceptance_rate,
                   "E_per_site": tdvp.E_local.real / N,
                   "kernel_ms_total": {k: v[0] for k, v in kernel_ms.items()},
                   "cg_iterations": tdvp.last_solve["iterations"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel": "mc_sweep",
                     "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
