"""ESS/s of one VMC step (sampling + E_loc + O_k + SR solve) on one MI355X:
learned block proposal (bf16 inference) vs the plain fp32 exchange sampler.

Workload: 2D J1 Heisenberg 10x10 periodic (N=100, 600 Pauli strings), complex
RBM with biases alpha=4 (PsiDeep, M=400, P=40,500, seeded complex Gaussian
sigma=0.01, final weights 1), 16,384 persistent chains in the S_z=0 sector,
angles on MFMA (bf16x3 split, fp32 accumulation), fp32 sweeps / E_loc / O_k,
fp64 reductions, SR by CG on the matrix-free S (fp64 vectors, from zero) to a
relative residual of 1e-2 (at most 100 iterations), the W parameters
preconditioned by Kronecker factors of S (spin covariance x hidden-unit
covariance, blocks of 64 units) and the biases by its diagonal, then the SR
update.  (The Jacobi-preconditioned CG, --solver cg, reaches only 3-5 % relative
residual in the same 100 iterations.)

Protocol (all on the GPU, untimed unless stated):
  1. the RBM is pre-optimised by 25 SR steps at lr 0.05 with the plain sampler
     (shift 1e-2, updates clipped to 0.02 per parameter), so that the
     autocorrelation is that of a structured state (E/N about -1.5), not of the
     near-uniform seed.  The run is kept short on purpose: the reference's layer-0
     activation is the series z^2/2 - z^4/12 + z^6/45, which grows without bound,
     and longer SR runs reach angles where log psi develops spikes (E_loc variance
     jumps by orders of magnitude, on some seeds within 30 steps); 25 steps were
     stable on every seed tried, and so were 10 further steps at the timed rate;
  2. training data: num_chains x 64 samples of the plain sampler; the proposal
     network (100 -> 256 -> 100, tanh; 2x4 / 4x2 plaquette blocks) is trained by
     the library's own forward / backward / Adam kernels;
  3. tau_int (Sokal window c = 5, batch means as a cross-check) of E_loc per sweep
     on 4096 chains x 400 sweeps after 300 plain thermalisation sweeps (the energy of
     the pre-optimised state relaxes over ~100 sweeps from a random start), for plain
     exchange sweeps and for learned block sweeps (the sweep count ceil(2 tau) is a
     step function of the estimate, so it is taken on enough chains not to flip with
     the noise);
  4. timed VMC steps of three paths on the same state:
       plain-config  : 100 exchange sweeps per sample (the workload's definition)
       plain-matched : ceil(2 tau_plain) exchange sweeps per sample
       learned       : ceil(2 tau_learned) block sweeps per sample (headline)
ESS per step = Ns * min(1, S_sw / (2 tau_int)).  Prints one JSON line.
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


def cpu_baseline(A, L, alpha, seed, seconds, sweeps):
    """Oracle (fp64 C restatement of the reference host path, OpenMP over chains)
    on a bounded sample of the same workload: chains x `sweeps` exchange sweeps,
    then E_loc and O_k per sample.  Excludes the SR solve (dense S of 40,500^2
    does not fit a CPU step), which makes the baseline optimistic."""
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
    chains = int(seconds / (sweeps * per_sweep * 1.2) * cores)
    chains = max(cores, min(4096, chains // cores * cores))
    t0 = time.perf_counter()
    r = orc.mc_run(o, chains, 1, sweeps, 0, move=orc.MOVE_EXCHANGE, seed=2, num_threads=cores)
    w = np.full(chains, 1.0 / chains)
    orc.tdvp_accumulate(op, o, r["samples"], w, with_S=False, num_threads=cores)
    dt = time.perf_counter() - t0
    return dict(samples_per_s=chains / dt, seconds=dt, chains=chains, cores=cores)


KERNELS = ("mc_sweep", "block_sweep", "theta", "local_energy", "o_k", "spin_sums", "sr_pack", "sr_ov", "sr_finish",
           "spin_gemm", "sr_y", "kfac_setup", "kfac_step")


class Stepper:
    """one VMC step: TDVP force on fresh samples, warm-started CG on (S + shift) x = F
    (K-FAC or Jacobi preconditioner), SR update clipped to `clip` per parameter (the
    polynomial activations diverge if a single step moves the weights too far)"""

    def __init__(self, A, psi, op, ens, lr, cg_iters, shift=1e-3, clip=0.02, solver="cg_kfac", rtol=1e-2, warm=True):
        self.A, self.psi, self.op, self.ens = A, psi, op, ens
        self.tdvp = A.TDVP(psi.num_params)
        self.lr, self.cg_iters, self.shift, self.clip = lr, cg_iters, shift, clip
        self.solver, self.rtol, self.warm = solver, rtol, warm
        self.x = None
        self.iters = []
        self.residuals = []
        # the host copy of the parameters is kept here (psi is only written by this
        # stepper while it runs), so a step does not read them back from the library
        self.params = psi.params.copy()

    def step(self):
        self.tdvp.eval_F(self.op, self.psi, self.ens)
        x = self.tdvp.solve(self.shift, self.solver, rtol=self.rtol, max_iter=self.cg_iters, raise_on_failure=False,
                            x0=self.x if self.warm else None)
        self.iters.append(self.tdvp.last_solve["iterations"])
        self.residuals.append(self.tdvp.last_solve["residual"])
        if not np.all(np.isfinite(x)):
            raise FloatingPointError("SR update is not finite")
        self.x = x
        d = self.lr * x
        m = np.abs(d).max()
        if m > self.clip:
            d *= self.clip / m
        self.params -= d
        self.psi.params = self.params

    def timed(self, steps, warmup):
        A = self.A
        for _ in range(warmup):
            self.step()
        A.profiler_reset()
        A.start_profiling()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        A.stop_profiling()    # synchronises the device
        dt = time.perf_counter() - t0
        kms = {k: A.profiler_query(k) for k in KERNELS}
        return dt / steps * 1e3, kms


THERM_SWEEPS = 300   # plain sweeps before a chain's first sample (also for learned paths)


def measure_tau(A, psi, op, proposal, chains=4096, T=400, seed=23):
    # 300 plain thermalisation sweeps: from a random S_z = 0 start the energy of this
    # state relaxes over ~100 sweeps (a series started after 30 sweeps is a transient).
    # The first pass over the persistent chains thermalises and is discarded; the series
    # and the acceptance rate (which counts every proposal of a pass, thermalisation
    # included, as the reference's does) come from the second pass.
    ens = A.MonteCarloSpins(chains * T, 1, THERM_SWEEPS, chains, move="exchange", seed=seed, persistent=True)
    if proposal is not None:
        ens.set_proposal(proposal, plain_sweeps=0)
    A.log_psi(psi, ens)
    el = A.local_energies(op, psi, ens).real.reshape(T, chains)
    return tau_int_sokal(el), tau_batch_means(el), ens.acceptance_rate, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=10)
    ap.add_argument("--alpha", type=int, default=4)
    ap.add_argument("--chains", type=int, default=16384)
    ap.add_argument("--sweeps", type=int, default=100)
    ap.add_argument("--opt-steps", type=int, default=25)
    ap.add_argument("--opt-lr", type=float, default=0.05)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--cg-iters", type=int, default=100)
    ap.add_argument("--solver", default="cg_kfac", choices=("cg_kfac", "cg"))
    ap.add_argument("--rtol", type=float, default=1e-2)
    ap.add_argument("--warm-solve", action="store_true",
                    help="start every CG solve from the previous step's solution (measured slower with K-FAC)")
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
    M = args.alpha * N

    # 1. pre-optimisation with the plain sampler (10 sweeps per sample), in fp64: the
    # state it reaches sets the measured workload (tau, sweeps per sample, CG
    # iterations), and a Monte Carlo trajectory flips accept decisions on perturbations
    # of the ratios at the fp32 rounding level, so an fp32 optimisation would land on
    # another state whenever an fp32 kernel changes its summation order
    t0 = time.perf_counter()
    psi.precision = "fp64"
    opt_ens = A.MonteCarloSpins(B, 10, 30, B, move="exchange", seed=5, persistent=True)
    # (the stability of this schedule was established with the Jacobi solve below)
    opt = Stepper(A, psi, op, opt_ens, args.opt_lr, args.cg_iters, shift=1e-2, solver="cg", rtol=1e-4)
    for _ in range(args.opt_steps):
        opt.step()
    psi.precision = "fp32"
    E_opt = opt.tdvp.E_local.real / N
    var_opt = opt.tdvp.var_H / N
    if not (var_opt < 50.0 and E_opt > -2.75):
        print(f"warning: pre-optimised state has E/N = {E_opt:.3f}, var/N = {var_opt:.3g} (outside the "
              f"physical range): autocorrelation estimates will not be meaningful", file=sys.stderr, flush=True)
    t_opt = time.perf_counter() - t0
    print(f"pre-optimised: E/N = {E_opt:.4f} after {args.opt_steps} SR steps ({t_opt:.1f} s)", file=sys.stderr,
          flush=True)

    # 2. training data from the plain sampler and proposal training
    t0 = time.perf_counter()
    data = A.MonteCarloSpins(B * 64, 2, 30, B, move="exchange", seed=7)
    A.log_psi(psi, data)
    blocks = A.rect_blocks(args.L, (2, 4))
    prop = A.BlockProposal(N, blocks, hidden=256, seed=11)
    loss0, _ = prop.loss_grad(data.last_samples()[:4096], blocks[np.arange(4096) % len(blocks)])
    A.profiler_reset()
    A.start_profiling()
    loss = prop.fit(data, steps=args.train_steps, batch=4096, lr=2e-3)
    A.stop_profiling()
    train_ms = A.profiler_query("proposal_train")[0]
    t_train = time.perf_counter() - t0
    print(f"proposal trained: loss {loss0:.4f} -> {loss:.4f} ({t_train:.1f} s)", file=sys.stderr, flush=True)

    # 3. autocorrelation of E_loc, one (plain or block) sweep apart
    tau_p, tau_p_bm, acc_p, el_p = measure_tau(A, psi, op, None)
    tau_l, tau_l_bm, acc_l, el_l = measure_tau(A, psi, op, prop)
    print(f"tau plain {tau_p:.2f} sweeps, tau learned {tau_l:.2f} block sweeps", file=sys.stderr, flush=True)

    # 4. timed paths
    paths = {}
    s_match = max(1, math.ceil(2 * tau_p))
    s_learn = max(1, math.ceil(2 * tau_l))
    for name, sweeps, proposal, tau in (("plain-config", args.sweeps, None, tau_p),
                                        ("plain-matched", s_match, None, tau_p),
                                        ("learned", s_learn, prop, tau_l)):
        ens = A.MonteCarloSpins(B, sweeps, THERM_SWEEPS, B, move="exchange", seed=17, persistent=True)
        if proposal is not None:
            ens.set_proposal(proposal, plain_sweeps=0)
        st = Stepper(A, psi.copy(), op, ens, args.lr, args.cg_iters, solver=args.solver, rtol=args.rtol,
                     warm=args.warm_solve)
        ms, kms = st.timed(args.steps, args.warmup)
        ess = B * min(1.0, sweeps / (2 * tau))
        paths[name] = dict(sweeps_per_sample=sweeps, ms_per_step=ms, ess_per_step=ess,
                           ess_per_s=ess / (ms * 1e-3), tau_int=tau, acceptance=ens.acceptance_rate,
                           E_per_site=st.tdvp.E_local.real / N, cg_iterations_mean=float(np.mean(st.iters)),
                           cg_residual_mean=float(np.mean(st.residuals)),
                           kernel_ms_total={k: v[0] for k, v in kms.items()},
                           kernel_launches={k: v[1] for k, v in kms.items()})
        print(f"{name}: {ms:.2f} ms/step, ESS/s {paths[name]['ess_per_s']:.0f}", file=sys.stderr, flush=True)
    head = paths["learned"]

    # energies of the two samplers on the same state (samples of step 3)
    def mean_se(el):
        cm = el.mean(axis=0)
        return float(cm.mean()), float(cm.std(ddof=1) / math.sqrt(el.shape[1]))
    Ep, sep = mean_se(el_p)
    El, sel = mean_se(el_l)

    # roofline of the headline path's dominant single-kernel launch class (each timer
    # below brackets exactly one kernel launch with hipEvents on the library stream;
    # sr_ov and spin_gemm are timed on every 16th launch, so that the events do not
    # stall the CG loop)
    def roofline(path, name):
        kt = paths[path]["kernel_ms_total"]
        launches = paths[path]["kernel_launches"][name]
        avg_ms = kt[name] / max(launches, 1)
        if name in ("spin_gemm", "sr_ov"):
            if name == "spin_gemm":
                # streams g (8 B per hidden unit), the fp32 coefficient and the transposed spin bits
                # (8 B per 64 samples and row, N + 1 rows padded to 16) of every sample once
                vpad = (N + 1 + 15) // 16 * 16
                work = B * (8.0 * M + 8.0 + 8.0 * vpad / 64)
            else:
                # streams g once, reads each configuration (8 B per word) and writes one fp32 complex
                # partial per column pair of 16 hidden units
                work = B * (8.0 * M + 8.0 * math.ceil(N / 64) + 8.0 * math.ceil(M / 16))
            r = {"bound": "hbm", "peak": 8000.0, "unit": "GB/s", "engine": "HBM3E"}
            achieved = work / (avg_ms * 1e-3) / 1e9
        else:
            if name == "block_sweep":
                bs = blocks.shape[1]
                moves = paths[path]["sweeps_per_sample"] * math.ceil(N / bs)
                # per block move: second layer 2 bs H, angle update of <= bs rows 4 bs M, log-cosh sum 26 M
                work = B * moves * (2.0 * bs * 256 + 4.0 * bs * M + 26.0 * M)
            elif name == "mc_sweep":
                work = B * paths[path]["sweeps_per_sample"] * N * 36.0 * M
            else:   # local_energy: 36 M per distinct off-diagonal configuration (~130 on the 10x10 AFM)
                work = B * 130 * 36.0 * M
            r = {"bound": "mfma", "peak": 157.3, "unit": "TFLOP/s",
                 "engine": "fp32 VALU (MI355X vector fp32 peak 157.3 TFLOP/s = fp32 matrix peak)"}
            achieved = work / (avg_ms * 1e-3) / 1e12
        traffic = measured_traffic(name) if r["bound"] == "hbm" else None
        return dict(r, achieved=achieved, frac=achieved / r["peak"], traffic=traffic, kernel=name,
                    avg_launch_ms=avg_ms, launches=launches, algorithmic_bytes=work)

    def measured_traffic(name):
        """HBM bytes per launch of the kernel behind timer `name` on this workload, from the
        committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when not measured)"""
        if (args.L, args.alpha, B) != (10, 4, 16384):
            return None
        path = os.path.join(ROOT, "profiles", "sr_matvec_heisenberg10x10_hbm_traffic.md")
        if not os.path.exists(path):
            return None
        for line in open(path):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) >= 9 and cols[1] == name:
                return float(cols[7]) * 1e6
        return None

    candidates = ("block_sweep", "sr_ov", "spin_gemm", "local_energy", "mc_sweep")
    dom = max(candidates, key=lambda k: head["kernel_ms_total"][k])
    roof = roofline("learned", dom)
    roof_plain = roofline("plain-config", "mc_sweep")

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(A, args.L, args.alpha, 3, args.cpu_seconds, args.sweeps)
        ess_frac = min(1.0, args.sweeps / (2 * tau_p))
        cpu = {"value": c["samples_per_s"] * ess_frac, "unit": "ESS/s", "cores": c["cores"], "kind": "port",
               "sample": f"{c['chains']} chains x {args.sweeps} exchange sweeps + E_loc + O_k in fp64 "
                         f"({c['seconds']:.1f} s, OpenMP over chains) on the seeded RBM; SR solve excluded; "
                         f"ESS fraction of the plain GPU path applied"}

    line = {
        "metric": "ESS/s with E_loc+O_k+SR, RBM N=100 alpha=4: learned-proposal bf16 vs plain fp32",
        "value": head["ess_per_s"], "unit": "ESS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"Heisenberg {args.L}x{args.L} PBC, complex RBM alpha={args.alpha} "
                               f"(P={psi.num_params}), {B} persistent chains, S_z=0; learned 2x4 block proposal "
                               f"(MLP {N}-256-{N}, bf16 inference) vs plain exchange sweeps; fp32 sweeps/E_loc/O_k, "
                               f"bf16x3 angle GEMM, {args.solver} SR to rtol {args.rtol} (fp64)",
                   "learned_vs_plain_config": head["ess_per_s"] / paths["plain-config"]["ess_per_s"],
                   "learned_vs_plain_matched": head["ess_per_s"] / paths["plain-matched"]["ess_per_s"],
                   "paths": paths,
                   "pre_optimisation": {"sr_steps": args.opt_steps, "lr": args.opt_lr, "E_per_site": E_opt,
                                        "var_per_site": var_opt, "seconds": t_opt},
                   "training": {"samples": B * 64, "adam_steps": args.train_steps, "batch": 4096,
                                "loss_before": loss0, "loss_after": loss, "device_ms": train_ms,
                                "seconds": t_train},
                   "tau": {"plain_sweeps": tau_p, "plain_batch_means": tau_p_bm,
                           "learned_block_sweeps": tau_l, "learned_batch_means": tau_l_bm,
                           "acceptance_plain": acc_p, "acceptance_learned": acc_l},
                   "energy_check": {"E_plain": Ep, "se_plain": sep, "E_learned": El, "se_learned": sel,
                                    "rel_diff": abs(El - Ep) / abs(Ep)},
                   "roofline_plain_sweep": roof_plain},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
