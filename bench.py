"""ESS/s of one VMC step (sampling + E_loc + O_k + SR solve) on one MI355X:
learned block proposal (bf16 inference) vs the plain fp32 samplers.

Workload: 2D J1 Heisenberg 10x10 periodic (N=100, 600 Pauli strings), complex
RBM with biases alpha=4 (PsiDeep, M=400, P=40,500, seeded complex Gaussian
sigma=0.01, final weights 1), 16,384 persistent chains, angles on MFMA (bf16x3
split, fp32 accumulation), fp32 sweeps / E_loc / O_k, fp64 reductions, SR by CG
on the matrix-free S (fp64 vectors, from zero) to the workload's relative residual
1e-6 (at most 200 iterations), the W parameters preconditioned by Kronecker
factors of S (spin covariance x hidden-unit covariance, blocks of 64 units) and the
biases by its diagonal, then the SR update.

Protocol (all on the GPU, untimed unless stated):
  1. the RBM is pre-optimised by 25 SR steps at lr 0.05 with the plain sampler
     (shift 1e-2, updates clipped to 0.02 per parameter), so that the
     autocorrelation is that of a structured state (E/N about -1.5), not of the
     near-uniform seed.  The run is kept short on purpose: the reference's layer-0
     activation is the series z^2/2 - z^4/12 + z^6/45, which grows without bound,
     and longer SR runs reach angles where log psi develops spikes;
  2. training data: num_chains x 64 samples of the plain sampler; the proposal
     network (100 -> 256 -> 100, tanh; 2x4 / 4x2 plaquette blocks) is trained by
     the library's own forward / backward / Adam kernels; a block move draws its
     pattern from the network's conditional law without the current pattern
     (Metropolized Gibbs, exact under the corrected Metropolis-Hastings ratio);
  3. tau_int (Sokal window c = 5, batch means as a cross-check) of E_loc per sweep
     on 4096 chains x 400 sweeps after 300 plain thermalisation sweeps, for plain
     exchange sweeps, plain single-flip sweeps (the full spin space) and learned
     block sweeps;
  4. timed VMC steps on the same state:
       plain-config       : 100 exchange sweeps per sample (the workload's definition)
       plain-matched      : ceil(2 tau) exchange sweeps per sample
       plain-flip-matched : ceil(2 tau) single-flip sweeps per sample (full space)
       learned            : ceil(2 tau) block sweeps per sample (headline)
       learned-rtol1e-2   : the learned path with the SR solve stopped at 1e-2
  5. energy check: both samplers on the same state (1,000 plain thermalisation
     sweeps) until the standard error of the mean (from independent chain means) is
     below 0.004 or 20 s have passed, and the fp32 device path
     against the fp64 device path on the same 4,096 samples;
  6. dense-S leg (1D Heisenberg ring N=64, M=256, P=16,704, 8,192 chains, 64
     exchange sweeps): S on fp64 matrix cores and the fp64 blocked Cholesky solve;
     CNN quench leg (TFIM 12x12, PsiCNN 2 x 3x3 x 8 channels, 16,384 single-flip
     chains, RK4 steps of dt = 0.005);
  7. CPU baseline: the fp64 oracle (sampling, E_loc, O_k, F and a matrix-free
     Jacobi CG to the same tolerance) on a bounded sample, at 1 core and all cores.
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


def cpu_baseline(A, L, alpha, seed, seconds, legs, tau_p):
    """The fp64 oracle (a C restatement of the reference host path, OpenMP over
    chains) timed on a bounded sample of the same workload: `chains` chains x `sweeps`
    exchange sweeps, then E_loc, O_k and F per sample and the SR solve (matrix-free
    Jacobi CG, shift 1e-3, relative residual 1e-6, at most 200 iterations).  legs:
    [(name, sweeps, threads)]; each is sized to about `seconds` from a calibration."""
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

    def step(chains, sweeps, threads):
        t0 = time.perf_counter()
        r = orc.mc_run(o, chains, 1, sweeps, 0, move=orc.MOVE_EXCHANGE, seed=2, num_threads=threads)
        w = np.full(chains, 1.0 / chains)
        acc = orc.tdvp_accumulate(op, o, r["samples"], w, with_S=False, num_threads=threads)
        x, it, res = orc.sr_cg_matfree(acc["O_samples"], w, acc["O"], acc["F"], 1e-3, rtol=1e-6, max_iter=200,
                                       num_threads=threads)
        return time.perf_counter() - t0, it

    out = {}
    for name, sweeps, threads in legs:
        cal = 4 * threads
        t, _ = step(cal, sweeps, threads)
        chains = int(seconds / max(t / cal, 1e-9))
        chains = max(2 * threads, min(4096, chains // threads * threads))
        dt, it = step(chains, sweeps, threads)
        ess = chains * min(1.0, sweeps / (2 * tau_p))
        out[name] = dict(ess_per_s=ess / dt, samples_per_s=chains / dt, seconds=dt, chains=chains, sweeps=sweeps,
                         threads=threads, cg_iterations=it)
    return out


KERNELS = ("mc_sweep", "block_sweep", "theta", "local_energy", "o_k", "spin_sums", "sr_pack", "sr_ov", "sr_finish",
           "spin_gemm", "sr_y", "kfac_setup", "kfac_step", "gram", "cholesky")


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
ENERGY_THERM_SWEEPS = 1000   # the same before the energy comparison (its transients must not bias 1e-4)


def measure_tau(A, psi, op, proposal, chains=4096, T=400, seed=23, move="exchange"):
    # 300 plain thermalisation sweeps: from a random start the energy of this state
    # relaxes over ~100 sweeps (a series started after 30 sweeps is a transient).
    # The first pass over the persistent chains thermalises and is discarded; the series
    # and the acceptance rate (which counts every proposal of a pass, thermalisation
    # included, as the reference's does) come from the second pass.
    ens = A.MonteCarloSpins(chains * T, 1, THERM_SWEEPS, chains, move=move, seed=seed, persistent=True)
    if proposal is not None:
        ens.set_proposal(proposal, plain_sweeps=0, exclude_current=True)
    A.log_psi(psi, ens)
    el = A.local_energies(op, psi, ens).real.reshape(T, chains)
    return tau_int_sokal(el), tau_batch_means(el), ens.acceptance_rate, el


def energy_run(A, psi, op, proposal, chains, target_se, max_seconds, seed, chunk=40, halves=None):
    """E on persistent chains (one plain or block sweep between samples) until the
    standard error of the mean from the independent chain means is below target_se or
    max_seconds have passed; returns (E, SE, samples, seconds, last chunk's configurations).
    With a dict as halves, E and SE of the first and of the second half of the run are
    stored in it (a mean that still moves between the halves is a chain that is still
    relaxing)."""
    ens = A.MonteCarloSpins(chains * chunk, 1, ENERGY_THERM_SWEEPS, chains, move="exchange", seed=seed,
                            persistent=True)
    if proposal is not None:
        ens.set_proposal(proposal, plain_sweeps=0, exclude_current=True)
    sums = np.zeros(chains)
    n = 0
    parts = []
    t0 = time.perf_counter()
    se = float("inf")
    while True:
        el = A.local_energies(op, psi, ens).real.reshape(chunk, chains)
        parts.append(el.sum(axis=0))
        sums += parts[-1]
        n += chunk
        cm = sums / n
        se = float(cm.std(ddof=1) / math.sqrt(chains))
        if se < target_se or time.perf_counter() - t0 > max_seconds:
            break
    if halves is not None and len(parts) >= 2:
        h = len(parts) // 2
        for name, sel in (("first", parts[:h]), ("second", parts[h:])):
            hm = np.sum(sel, axis=0) / (len(sel) * chunk)
            halves[name] = {"E": float(hm.mean()), "se": float(hm.std(ddof=1) / math.sqrt(chains)),
                            "samples_per_chain": len(sel) * chunk}
    return float(cm.mean()), se, n * chains, time.perf_counter() - t0, ens.last_samples()[:4096]


def fp64_check(A, psi, op, confs):
    """the fp32 device path against the fp64 device path on the same samples"""
    out = {}
    res = {}
    for prec in ("fp32", "fp64"):
        q = psi.copy()
        q.precision = prec
        t = A.TDVP(q.num_params)
        t.eval_F(op, q, A.SampleList(q.num_sites, confs))
        res[prec] = (t.E_local_samples, t.E_local, t.F_vector, t.O_k_vector)
    e32, e64 = res["fp32"], res["fp64"]
    out["E_loc_max_rel"] = float((np.abs(e32[0] - e64[0]) / np.abs(e64[0])).max())
    out["E_rel"] = abs(e32[1] - e64[1]) / abs(e64[1])
    out["F_rel_norm"] = float(np.linalg.norm(e32[2] - e64[2]) / np.linalg.norm(e64[2]))
    out["O_rel_norm"] = float(np.linalg.norm(e32[3] - e64[3]) / np.linalg.norm(e64[3]))
    out["samples"] = int(len(confs))
    return out


def dense_S_leg(A, steps=2, warmup=1):
    """1D Heisenberg ring N=64, PsiDeep M=256 (P=16,704), 8,192 persistent exchange chains,
    64 sweeps per sample; S on fp64 matrix cores, fp64 blocked Cholesky of S + 1e-3"""
    N, M, B = 64, 256, 8192
    psi = A.seeded_rbm_deep(N, M, seed=2, sigma=0.01, final_weight=1.0)
    psi.precision = "fp32"
    op = A.Operator(A.hamiltonians.heisenberg_chain(N))
    ens = A.MonteCarloSpins(B, 64, 64, B, move="exchange", seed=3, persistent=True)
    tdvp = A.TDVP(psi.num_params)
    P = psi.num_params
    params = psi.params.copy()

    def one():
        tdvp.eval(op, psi, ens)
        x = tdvp.solve(1e-3, "cholesky")
        d = 0.02 * x
        m = np.abs(d).max()
        if m > 0.02:
            d *= 0.02 / m
        params[:] -= d
        psi.params = params
        return tdvp.last_solve["residual"]

    for _ in range(warmup):
        one()
    A.profiler_reset()
    A.start_profiling()
    t0 = time.perf_counter()
    res = [one() for _ in range(steps)]
    A.stop_profiling()
    ms = (time.perf_counter() - t0) / steps * 1e3
    g_ms, g_n = A.profiler_query("gram")
    c_ms, c_n = A.profiler_query("cholesky")
    gram_flop = 4.0 * B * P * P            # upper-triangular half of the Hermitian O^H W O (real flops)
    chol_flop = 4.0 / 3.0 * P ** 3         # complex Cholesky (real flops); the two triangular solves excluded
    g_avg, c_avg = g_ms / max(g_n, 1), c_ms / max(c_n, 1)
    return {"workload": "Heisenberg ring N=64 PBC, PsiDeep M=256 (P=16,704), 8192 persistent exchange chains, "
                        "64 sweeps per sample, fp32 E_loc/O_k, dense S (fp64 matrix cores) + fp64 Cholesky, "
                        "shift 1e-3",
            "ms_per_step": ms, "steps": steps, "solve_residual_max": float(max(res)),
            "gram": {"ms": g_avg, "algorithmic_flop": gram_flop, "achieved_tflops": gram_flop / g_avg * 1e-9,
                     "peak_tflops": 78.6, "frac": gram_flop / g_avg * 1e-9 / 78.6, "engine": "fp64 MFMA"},
            "cholesky": {"ms": c_avg, "algorithmic_flop": chol_flop, "achieved_tflops": chol_flop / c_avg * 1e-9,
                         "peak_tflops": 78.6, "frac": chol_flop / c_avg * 1e-9 / 78.6, "engine": "fp64 MFMA"}}


def cnn_quench_leg(A, steps=2, warmup=1, progress=False, precision="fp32", shift=1e-3):
    """TFIM 12x12 (h = 3.04) quench from the near-uniform state: PsiCNN with 2 layers of 3x3
    periodic kernels, 8 channels (648 parameters, weights complex-Gaussian 1e-3), 16,384
    persistent single-flip chains (one sweep between samples), RK4 with dt = 0.005 and the
    fp64 Cholesky solve of S + 1e-3 on every right-hand side"""
    L = 12
    N = L * L
    rng = np.random.default_rng(4)
    n = 1 * 8 * 9 + 8 * 8 * 9
    params = 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    psi = A.PsiCNN([1, L, L], [8, 8], [[1, 3, 3], [1, 3, 3]], np.zeros(N, dtype=np.uint32), params, 1.0, 0.0)
    psi.precision = precision
    op = A.Operator(A.hamiltonians.transverse_ising_square(L, 1.0, 3.04))
    B = 16384
    ens = A.MonteCarloSpins(B, 1, 50, B, move="flip", seed=3, persistent=True)
    rk = A.RK4(psi, op, ens, 0.005, shift=shift, method="cholesky")
    for _ in range(warmup):
        rk.step()
    A.profiler_reset()
    A.start_profiling()
    t0 = time.perf_counter()
    failure = None
    pnorm = [float(np.linalg.norm(psi.params))]
    for k in range(steps):
        try:
            rk.step()
            pnorm.append(float(np.linalg.norm(psi.params)))
        except RuntimeError as e:
            # the record says where the evolution stopped and what S looked like there
            S = rk.tdvp.S_matrix
            ev = np.linalg.eigvalsh(S) if np.all(np.isfinite(S)) else None
            El = rk.tdvp.E_local_samples
            failure = {"step": k, "t": rk.t, "error": str(e), "S_finite": bool(np.all(np.isfinite(S))),
                       "params_finite": bool(np.all(np.isfinite(psi.params))),
                       "E_loc_samples_nonfinite": int(np.sum(~np.isfinite(El))),
                       "S_diag_min": float(np.min(S.diagonal().real)), "S_diag_max": float(np.max(S.diagonal().real)),
                       "S_eig_min": float(ev[0]) if ev is not None else None,
                       "S_eig_max": float(ev[-1]) if ev is not None else None,
                       "E_last": complex(rk.tdvp.E_local).__repr__()}
            print(f"quench stopped at step {k}: {failure}", file=sys.stderr, flush=True)
            steps = max(k, 1)
            break
        if progress and ((k + 1) % 20 == 0 or k + 1 > 100 and (k + 1) % 2 == 0):
            print(f"quench step {k + 1}: t = {rk.t:.3f}, E = {rk.energies[-4].real:.4f}, |params| = {pnorm[-1]:.4g}, "
                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    A.stop_profiling()
    ms = (time.perf_counter() - t0) / steps * 1e3
    kms = {k: A.profiler_query(k) for k in ("mc_sweep", "local_energy", "o_k", "spin_sums", "gram", "cholesky",
                                            "log_psi")}
    sweep_ms, sweep_n = kms["mc_sweep"]
    # work model of one single-flip proposal with local updates (a flip away from the
    # lattice edge changes 9 layer-0 and 25 layer-1 sites): complex MACs 9*8*9 + 25*8*8*9,
    # 8 real flops each
    work = B * N * 8.0 * (9 * 8 * 9 + 25 * 8 * 8 * 9) * max(sweep_n, 1)
    E = np.array(rk.energies).real
    # the energy at the first evaluation of every step (4 evaluations per RK4 step): a
    # real-time evolution conserves it up to the Monte Carlo noise and the shift
    E_steps = E[::4]
    return {"workload": "TFIM 12x12 PBC h=3.04 quench, PsiCNN 2 x (3x3, 8 channels), P=648, 16384 persistent "
                        "single-flip chains, 1 sweep per sample, RK4 dt=0.005, fp32 sampling/E_loc/O_k, dense S "
                        "+ fp64 Cholesky (shift 1e-3)",
            "ms_per_rk4_step": ms, "steps": steps, "tdvp_evaluations_per_step": 4,
            "kernel_ms_per_step": {k: v[0] / steps for k, v in kms.items()},
            "kernel_launches": {k: v[1] for k, v in kms.items()},
            "energy_first_last": [float(E[0]), float(E[-1])],
            "energy_per_step": [float(x) for x in E_steps],
            "energy_max_rel_change": float(np.max(np.abs(E_steps - E_steps[0])) / abs(E_steps[0])),
            "failure": failure, "precision": precision, "shift": shift,
            "param_norm_per_step": pnorm,
            "sweep_roofline": {"bound": "valu", "kernel": "cnn_mc_inc (local updates)",
                               "achieved_tflops": work / (sweep_ms * 1e-3) / 1e12 if sweep_ms else None,
                               "peak_tflops": 157.3,
                               "frac": work / (sweep_ms * 1e-3) / 1e12 / 157.3 if sweep_ms else None}}


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
    ap.add_argument("--opt-therm", type=int, default=30,
                    help="thermalisation sweeps of the pre-optimisation's persistent chains")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--cg-iters", type=int, default=200)
    ap.add_argument("--solver", default="cg_kfac", choices=("cg_kfac", "cg"))
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--warm-solve", action="store_true",
                    help="start every CG solve from the previous step's solution")
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--energy-se", type=float, default=0.004)
    ap.add_argument("--energy-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the energy check and the dense-S leg")
    ap.add_argument("--quench-only", type=int, default=0, metavar="STEPS",
                    help="run only the 12x12 CNN quench leg for STEPS RK4 steps and print its record")
    ap.add_argument("--quench-precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--quench-shift", type=float, default=1e-3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import annongpu_amd as A
    A.setDevice(0)
    if args.quench_only > 0:
        print(json.dumps(cnn_quench_leg(A, steps=args.quench_only, progress=True, precision=args.quench_precision,
                                        shift=args.quench_shift)))
        return
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
    opt_ens = A.MonteCarloSpins(B, 10, args.opt_therm, B, move="exchange", seed=5, persistent=True)
    opt = Stepper(A, psi, op, opt_ens, args.opt_lr, 100, shift=1e-2, solver="cg", rtol=1e-4)
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
    del data
    print(f"proposal trained: loss {loss0:.4f} -> {loss:.4f} ({t_train:.1f} s)", file=sys.stderr, flush=True)

    # 3. autocorrelation of E_loc, one (plain or block) sweep apart
    tau_p, tau_p_bm, acc_p, _ = measure_tau(A, psi, op, None)
    tau_f, tau_f_bm, acc_f, _ = measure_tau(A, psi, op, None, move="flip")
    tau_l, tau_l_bm, acc_l, _ = measure_tau(A, psi, op, prop)
    print(f"tau: plain exchange {tau_p:.2f}, plain single-flip {tau_f:.2f} sweeps, learned {tau_l:.2f} block sweeps",
          file=sys.stderr, flush=True)

    # 4. timed paths
    paths = {}
    s_match = max(1, math.ceil(2 * tau_p))
    s_flip = max(1, math.ceil(2 * tau_f))
    s_learn = max(1, math.ceil(2 * tau_l))
    for name, sweeps, proposal, tau, move, rtol, iters in (
            ("plain-config", args.sweeps, None, tau_p, "exchange", args.rtol, args.cg_iters),
            ("plain-matched", s_match, None, tau_p, "exchange", args.rtol, args.cg_iters),
            ("plain-flip-matched", s_flip, None, tau_f, "flip", args.rtol, args.cg_iters),
            ("learned", s_learn, prop, tau_l, "exchange", args.rtol, args.cg_iters),
            ("learned-rtol1e-2", s_learn, prop, tau_l, "exchange", 1e-2, 100)):
        ens = A.MonteCarloSpins(B, sweeps, THERM_SWEEPS, B, move=move, seed=17, persistent=True)
        if proposal is not None:
            ens.set_proposal(proposal, plain_sweeps=0, exclude_current=True)
        st = Stepper(A, psi.copy(), op, ens, args.lr, iters, solver=args.solver, rtol=rtol, warm=args.warm_solve)
        ms, kms = st.timed(args.steps, args.warmup)
        ess = B * min(1.0, sweeps / (2 * tau))
        paths[name] = dict(sweeps_per_sample=sweeps, move=move if proposal is None else "learned block",
                           sr_rtol=rtol, sr_max_iter=iters, ms_per_step=ms, ess_per_step=ess,
                           ess_per_s=ess / (ms * 1e-3), tau_int=tau, acceptance=ens.acceptance_rate,
                           E_per_site=st.tdvp.E_local.real / N, cg_iterations_mean=float(np.mean(st.iters)),
                           cg_residual_mean=float(np.mean(st.residuals)),
                           cg_residual_max=float(np.max(st.residuals)),
                           kernel_ms_total={k: v[0] for k, v in kms.items()},
                           kernel_launches={k: v[1] for k, v in kms.items()})
        if proposal is not None:
            paths[name]["block_stats_last_step"] = ens.block_stats
        print(f"{name}: {ms:.2f} ms/step, ESS/s {paths[name]['ess_per_s']:.0f}, CG {np.mean(st.iters):.0f} it "
              f"to {np.mean(st.residuals):.1e}", file=sys.stderr, flush=True)
    head = paths["learned"]

    # 5. energies of the two samplers on the same state, and fp32 against fp64
    energy = None
    dense = None
    cnn = None
    if not args.no_extra:
        hp, hl = {}, {}
        Ep, sep, n_p, t_p, confs = energy_run(A, psi, op, None, B, args.energy_se, args.energy_seconds, 41,
                                              halves=hp)
        El, sel, n_l, t_l, _ = energy_run(A, psi, op, prop, B, args.energy_se, args.energy_seconds, 43, halves=hl)
        energy = {"E_plain": Ep, "se_plain": sep, "samples_plain": n_p, "E_learned": El, "se_learned": sel,
                  "samples_learned": n_l, "rel_diff": abs(El - Ep) / abs(Ep),
                  "rel_se_combined": math.hypot(sep, sel) / abs(Ep),
                  "halves_plain": hp, "halves_learned": hl,
                  "fp32_vs_fp64_same_samples": fp64_check(A, psi, op, confs)}
        print(f"energy: plain {Ep:.4f} +- {sep:.4f}, learned {El:.4f} +- {sel:.4f}", file=sys.stderr, flush=True)
        for name, hv in (("plain", hp), ("learned", hl)):
            if hv:
                print(f"energy halves ({name}): {hv['first']['E']:.4f} +- {hv['first']['se']:.4f} then "
                      f"{hv['second']['E']:.4f} +- {hv['second']['se']:.4f}", file=sys.stderr, flush=True)
        # 6. dense S and Cholesky
        dense = dense_S_leg(A)
        cnn = cnn_quench_leg(A)
        print(f"cnn quench leg: {cnn['ms_per_rk4_step']:.1f} ms per RK4 step", file=sys.stderr, flush=True)
        print(f"dense S leg: {dense['ms_per_step']:.1f} ms/step, gram {dense['gram']['ms']:.1f} ms "
              f"({dense['gram']['achieved_tflops']:.1f} TFLOP/s), cholesky {dense['cholesky']['ms']:.1f} ms "
              f"({dense['cholesky']['achieved_tflops']:.1f} TFLOP/s)", file=sys.stderr, flush=True)

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
                # per proposal: second layer 2 bs H; per proposal that is not a redraw of the
                # current pattern: log-cosh sum 26 M; per flipped row entering a ratio: 4 M
                # (device counters of the last timed step, scaled to one launch)
                bs = blocks.shape[1]
                st = paths[path]["block_stats_last_step"]
                work = st["proposals"] * 2.0 * bs * 256 + (st["proposals"] - st["null_moves"]) * 26.0 * M \
                    + st["rows_evaluated"] * 4.0 * M
            elif name == "mc_sweep":
                work = B * paths[path]["sweeps_per_sample"] * N * 36.0 * M
            else:   # local_energy: 36 M per distinct off-diagonal configuration (~130 on the 10x10 AFM)
                work = B * 130 * 36.0 * M
            r = {"bound": "valu", "peak": 157.3, "unit": "TFLOP/s", "engine": "fp32 VALU (MI355X vector fp32 peak)"}
            achieved = work / (avg_ms * 1e-3) / 1e12
        traffic = measured_traffic(name) if r["bound"] == "hbm" else None
        return dict(r, achieved=achieved, frac=achieved / r["peak"], traffic=traffic, kernel=name,
                    avg_launch_ms=avg_ms, launches=launches, algorithmic_work_per_launch=work)

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
    extra_roofs = {k: roofline("learned", k) for k in candidates if k != dom and head["kernel_launches"][k] > 0}
    roof_plain = roofline("plain-config", "mc_sweep")

    cpu = None
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        legs = [("all_cores_matched", s_match, cores), ("one_core_matched", s_match, 1),
                ("all_cores_config", args.sweeps, cores)]
        c = cpu_baseline(A, args.L, args.alpha, 3, args.cpu_seconds, legs, tau_p)
        m = c["all_cores_matched"]
        cpu = {"value": m["ess_per_s"], "unit": "ESS/s", "cores": cores, "kind": "port",
               "sample": f"{m['chains']} chains x {s_match} exchange sweeps + E_loc + O_k + F in fp64 + matrix-free "
                         f"Jacobi CG (shift 1e-3, rtol 1e-6, <= 200 it: {m['cg_iterations']}) on the seeded RBM "
                         f"({m['seconds']:.1f} s, OpenMP over chains, {cores} threads); ESS fraction from the "
                         f"plain GPU tau",
               "legs": c}

    line = {
        "metric": "ESS/s with E_loc+O_k+SR, RBM N=100 alpha=4: learned-proposal bf16 vs plain fp32",
        "value": head["ess_per_s"], "unit": "ESS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"Heisenberg {args.L}x{args.L} PBC, complex RBM alpha={args.alpha} "
                               f"(P={psi.num_params}), {B} persistent chains, S_z=0; learned 2x4 block proposal "
                               f"(MLP {N}-256-{N}, bf16 inference); fp32 sweeps/E_loc/O_k, bf16x3 angle GEMM, "
                               f"{args.solver} SR to rtol {args.rtol}, <= {args.cg_iters} iterations (fp64)",
                   "learned_vs_plain_matched": head["ess_per_s"] / paths["plain-matched"]["ess_per_s"],
                   "learned_vs_plain_flip_matched": head["ess_per_s"] / paths["plain-flip-matched"]["ess_per_s"],
                   "learned_vs_plain_config": head["ess_per_s"] / paths["plain-config"]["ess_per_s"],
                   "paths": paths,
                   "pre_optimisation": {"sr_steps": args.opt_steps, "lr": args.opt_lr, "therm_sweeps": args.opt_therm, "E_per_site": E_opt,
                                        "var_per_site": var_opt, "seconds": t_opt},
                   "training": {"samples": B * 64, "adam_steps": args.train_steps, "batch": 4096,
                                "loss_before": loss0, "loss_after": loss, "device_ms": train_ms,
                                "seconds": t_train},
                   "tau": {"plain_exchange_sweeps": tau_p, "plain_exchange_batch_means": tau_p_bm,
                           "plain_flip_sweeps": tau_f, "plain_flip_batch_means": tau_f_bm,
                           "learned_block_sweeps": tau_l, "learned_batch_means": tau_l_bm,
                           "acceptance_plain_exchange": acc_p, "acceptance_plain_flip": acc_f,
                           "acceptance_learned": acc_l},
                   "energy_check": energy,
                   "dense_S": dense,
                   "cnn_quench": cnn,
                   "roofline_other_kernels": extra_roofs,
                   "roofline_plain_sweep": roof_plain},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
