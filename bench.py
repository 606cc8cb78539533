#!/usr/bin/env python3
"""Benchmark of the GIP fractional-step Navier–Stokes hot path on one MI355X.

Workload (the mixed-precision configuration, 256^3): the uniform 256^3
lid-driven-cavity pressure Laplacian (7-point, 117,440,512 nnz, pinned),
right-hand side a seeded N(0,1) field projected to zero mean.  A "step" is
one Jacobi-PCG solve of that system to a 1e-10 FP64 relative true residual
||b - A x||_2/||b||_2 with the mixed-precision solver (FP32 operator,
vectors and SpMV inside the FP64 defect-correction loop); `value` is its
time in ms.  Reported beside it: the same solve with the FP64 solver, the
FP64 SpMV HBM rate on the Laplacian (`roofline`), the FP32 product of the
timed mixed solves (the solve's largest kernel, timed inside the solves:
`roofline_solve`), the time-step rate of full Navier–Stokes steps at 256^3 (Re=1000, PCG to 1e-10,
FP64 vs mixed), the reference CPU path's PCG on the same system
(`cpu_baseline`), and ten steps of the 64-cavity batch beside the same
instances stepped one after another (`batch`).  Inputs are resident in HBM
before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--ns-steps 20]

Other suites (one JSON line each, not the default):
    --suite batch     64 uneven cavities (n ~ U{32..96}, Re log-uniform in
                      [100, 3200]) advanced 50 steps as one batch vs one by one
    --suite adjoint   64^3, Re=400, 200 steps, checkpoints every 20: forward +
                      backward time and 16 central finite-difference checks
    --suite reference the reference's own integrator (GIP_DNS over the Multicore
                      specification, one rank, one core, -O2) beside the GPU
                      integrator on the same cases (C100K and generated 46^3 /
                      96^3 cavities, fixed 100 PCG iterations): ms per step
    --suite clustered 128^3 tanh-clustered mesh (beta 1.5), Re=1000, 100 steps,
                      PCG to 1e-10: steps/s FP64 vs mixed, the L SpMV rate and
                      the FP64/mixed PCG on a zero-mean N(0,1) right-hand side
    --suite ns256     the 256^3 Navier-Stokes steps (Re=1000, pressure to 1e-10)
                      in mixed precision and FP64 on the GPU against the FP64
                      oracle, error per field (default 2 steps)
    --suite cpu       the CPU-baseline plan: the reference integrator at 1 and 4
                      ranks, as shipped (-O0) and -O3, on the generated 46^3
                      cavity and C100K, beside the GPU with 1 and 4 partitions
    --probe spmv      a few launches of the FP64 SpMV, eight iterations of each
                      PCG, and a plain copy of known size, for rocprofv3 --pmc
                      traffic runs
    --probe barrier   the cost of a device-wide barrier (cooperative launch)
                      among 8 to 1024 workgroups, relaxed and fenced
    --probe stream    the practical HBM ceiling: copy, read, write and the PCG
                      direction sweep's mix (3 reads, 2 writes) over 2 GiB
                      arrays, in 8- and 16-byte, plain and nontemporal
                      accesses, one access per thread or grid-strided
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import shutil
import subprocess
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "PCG ms to 1e-10 resid + steps/s, FP64 vs mixed, 256³ cavity; SpMV HBM GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(nnz, rows, cols):
    """Algorithmic bytes of one FP64 CSR SpMV: values+indices, row pointers,
    the output and one read of every input entry."""
    return 12 * nnz + 4 * (rows + 1) + 8 * rows + 8 * cols


def time_spmv(hip, dev, A, x, y, reps):
    e0, e1 = hip.Event(dev), hip.Event(dev)
    for _ in range(3):
        A.spmv(x, y)
    e0.record()
    for _ in range(reps):
        A.spmv(x, y)
    e1.record()
    return e0.elapsed_ms(e1) / reps


def true_relres(hip, dev, A, b_dev, x_dev, n, n_all, bnorm):
    ax = dev.array(n_all)
    hip.fill(ax, 0.0)
    A.spmv(x_dev, ax)
    hip.axpby(1.0, b_dev, -1.0, ax, n)
    return np.sqrt(hip.dot(ax, ax, n)) / bnorm


def pcg_leg(hip, dev, A, b_dev, n, n_all, precision, steps, warmup, maxit, bnorm):
    solver = hip.PCG(A, n_all)
    x = dev.array(n_all)
    z = dev.array(n_all)
    hip.fill(z, 0.0)
    e0, e1 = hip.Event(dev), hip.Event(dev)
    times, prod, res = [], [], None
    for k in range(warmup + steps):
        # time_products: one product launch per host poll (every 32
        # iterations) between two events on the solver's stream
        e0.record()
        res = solver.solve(b_dev, x, z, maxit=maxit, tol=1e-10, precision=precision, inner_tol=0.1, poll=32,
                           time_products=True)
        e1.record()
        ms = e0.elapsed_ms(e1)
        if k >= warmup:
            times.append(ms)
            prod.append(res.product_ms)
        log(f"  pcg {'mixed' if precision else 'fp64'} solve {k}: {ms:.1f} ms, {res.iters} it, "
            f"{res.corrections} corr, status {res.status}")
    rel = true_relres(hip, dev, A, b_dev, x, n, n_all, bnorm)
    solver.destroy()
    return {"ms": float(np.mean(times)), "ms_min": float(np.min(times)), "iters": res.iters,
            "corrections": res.corrections, "true_relres": float(rel), "status": res.status,
            "product_us": 1e3 * float(np.median(prod)), "product_samples_per_solve": res.product_samples}


def ns_leg(dns, cav, mode, steps, warmup):
    sim = dns.Simulation(cavity=cav, gamma=1e-3, pcg_mode=mode, pcg_tol=1e-10, pcg_maxit=20000)
    for _ in range(warmup):
        sim.step()
    sim.sync()
    t0 = time.perf_counter()
    its = []
    phases = {k: 0.0 for k in sim.PHASES}
    for k in range(steps):
        s = sim.step()
        its.append(s["pcg_iters"])
        # device time per phase (events at the phase boundaries; read after
        # the step, whose CFL phase has already synchronised)
        for name, ms in sim.phase_ms().items():
            phases[name] += ms / steps
    sim.sync()
    dt = time.perf_counter() - t0
    fields = {k: sim.field(k) for k in ("u", "v", "w", "p")}
    sim.close()
    return {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps, "pcg_iters_mean": float(np.mean(its)),
            "phase_ms": {k: round(v, 3) for k, v in phases.items()}}, fields


def spmv32_bytes(nnz, rows, cols):
    """Algorithmic bytes of one FP32-value SpMV (int32 indices)."""
    return 8 * nnz + 4 * (rows + 1) + 4 * rows + 4 * cols


def measured_traffic(n, kernel):
    """HBM bytes per launch of a kernel of the n^3 PCG workload, from the
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy kernel of known size): its `traffic_bytes_per_launch[kernel]` line."""
    f = ROOT / "profiles" / f"pcg_kernels_{n}_traffic.txt"
    if not f.exists():
        return None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    for line in f.read_text().splitlines():
        if line.startswith(key):
            return int(float(line.split(":", 1)[1]))
    return None


def stream_ceiling():
    """The best measured streaming rates (GB/s) from the committed
    `--probe stream` summary, or None."""
    f = ROOT / "profiles" / "stream_ceiling.txt"
    if not f.exists():
        return None
    out = {}
    for line in f.read_text().splitlines():
        if line.startswith("best_gbs["):
            k, v = line[len("best_gbs["):].split("]:", 1)
            out[k] = float(v)
    return out or None


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "cpu_count": os.cpu_count(), "usable_cores": usable}


def cpu_baseline(n, iters, seed, fp64_iters):
    """The reference's own GIP_LinearSolver + Multicore spec (oracle/_ref, the
    reference sources built -O3 -march=x86-64-v3, one rank, one core) on the
    same Laplacian and RHS, for a bounded number of fixed iterations, then
    extrapolated (labelled so) to the FP64 solve's iteration count."""
    exe = ROOT / "oracle" / "_ref" / "gip_ref_harness_O3"
    if not exe.exists():
        return None
    out = subprocess.run([str(exe), "pcg", str(n), str(iters), str(seed)], capture_output=True, text=True,
                         timeout=600, env=dict(os.environ, OMP_NUM_THREADS="1"))
    if out.returncode != 0:
        log("cpu baseline failed:", out.stderr[-500:])
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["ms_per_iter"] * fp64_iters, "unit": "ms", "cores": 1, "kind": "reference",
            "ms_per_iter": r["ms_per_iter"], "extrapolated": True, "measured_iters": iters,
            "measured_ms": r["solve_ms"],
            "sample": f"reference GIP_LinearSolver::solve (Multicore spec, 1 rank, -O3 -march=x86-64-v3) on the "
                      f"{n}^3 Laplacian: {iters} fixed Jacobi-PCG iterations timed ({r['solve_ms']:.0f} ms); value "
                      f"EXTRAPOLATED to the {fp64_iters} iterations the FP64 solve needs to 1e-10",
            "host": host_cpu()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ns-steps", type=int, default=None,
                    help="Navier-Stokes steps (default 20 in the default run, 2 in --suite ns256)")
    ap.add_argument("--ns-n", type=int, default=None)
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--skip-ns", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true", help="default run: leave out the 64-cavity batch leg")
    ap.add_argument("--suite", choices=("pcg", "batch", "adjoint", "clustered", "reference", "cpu", "ns256"),
                    default="pcg")
    ap.add_argument("--batch-eager", action="store_true",
                    help="--suite batch: launch every kernel from the host instead of replaying captured graphs")
    ap.add_argument("--skip-sequential", action="store_true",
                    help="--suite batch: time the batch only, not the instances one after another")
    ap.add_argument("--probe", choices=("spmv", "stream", "barrier"), default=None)
    ap.add_argument("--summarize", choices=("traffic", "stream", "kernels", "barrier"), default=None,
                    help="write a profiles/ summary from rocprofv3 --pmc outputs or a --probe stream JSON line")
    ap.add_argument("--fetch", default=None, help="--summarize traffic: the FETCH_SIZE run's output directory")
    ap.add_argument("--write", default=None, help="--summarize traffic: the WRITE_SIZE run's output directory")
    ap.add_argument("--json", default=None, help="--summarize stream/barrier: the probe's JSON output")
    ap.add_argument("--db", default=None, help="--summarize kernels: a rocprofv3 --kernel-trace results database")
    ap.add_argument("--title", default="", help="--summarize kernels: the workload and the command that profiled it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every partition of a case runs on one device "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)
    if args.summarize == "traffic":
        return summary_traffic(args)
    if args.summarize == "stream":
        return summary_stream(args)
    if args.summarize == "kernels":
        return summary_kernels(args)
    if args.summarize == "barrier":
        return summary_barrier(args)
    if args.probe == "stream":
        return probe_stream(args)
    if args.probe == "barrier":
        return probe_barrier(args)
    if args.probe:
        return probe_spmv(args)
    if args.suite == "batch":
        return batch_suite(args)
    if args.suite == "adjoint":
        return adjoint_suite(args)
    if args.suite == "clustered":
        return clustered_suite(args)
    if args.suite == "reference":
        return reference_suite(args)
    if args.suite == "cpu":
        return cpu_suite(args)
    if args.suite == "ns256":
        return ns256_suite(args)

    from gip_amd import cavity, dns, hip

    n = args.n
    t0 = time.perf_counter()
    cav = cavity.generate(n, gamma=1e-3, ops=cavity.OP_BITS["L"])
    L = cav.op("L")
    N, NA = cav.n_inner, cav.n_all
    b = np.zeros(NA)
    b[:N] = cavity.rhs_normal(N, args.seed)
    bnorm = float(np.linalg.norm(b[:N]))
    log(f"generated {n}^3 Laplacian: {L.nnz} nnz in {time.perf_counter() - t0:.1f} s")

    dev = hip.Device(0)
    A = hip.Matrix.from_csr(dev, L, fp32=True)
    info = A.info()
    b_dev = dev.to_device(b)
    del cav, L

    # FP64 SpMV on the Laplacian (dominant kernel)
    x = dev.to_device(np.random.default_rng(1).standard_normal(NA))
    y = dev.array(NA)
    spmv_ms = time_spmv(hip, dev, A, x, y, 50)
    alg = spmv_bytes(info["nnz"], N, NA)
    spmv_gbs = alg / (spmv_ms * 1e-3) / 1e9
    log(f"spmv fp64: {spmv_ms:.4f} ms, {spmv_gbs:.0f} GB/s algorithmic ({alg / 1e9:.3f} GB)")
    # the FP32 product of the mixed solver's inner iteration
    x32 = dev.to_device(np.random.default_rng(2).standard_normal(NA).astype(np.float32))
    y32 = dev.array(NA, np.float32)
    e0, e1 = hip.Event(dev), hip.Event(dev)
    for _ in range(3):
        A.spmv32(x32, y32)
    e0.record()
    for _ in range(50):
        A.spmv32(x32, y32)
    e1.record()
    spmv32_ms = e0.elapsed_ms(e1) / 50
    alg32 = spmv32_bytes(info["nnz"], N, NA)
    log(f"spmv fp32: {spmv32_ms:.4f} ms, {alg32 / (spmv32_ms * 1e-3) / 1e9:.0f} GB/s algorithmic ({alg32 / 1e9:.3f} GB)")
    del x, y, x32, y32

    fp64 = pcg_leg(hip, dev, A, b_dev, N, NA, hip.PCG_FP64, args.steps, args.warmup, 20000, bnorm)
    mixed = pcg_leg(hip, dev, A, b_dev, N, NA, hip.PCG_MIXED, args.steps, args.warmup, 20000, bnorm)
    it_fp64 = alg + 88 * N  # FP64 PCG iteration: SpMV + x, r, p, Ap, Dinv, s streams
    # FP32 inner iteration: SpMV + the same streams in FP32 (the deflation's
    # A1 comes from its list of nonzeros, one for this pinned Laplacian, and
    # is not streamed)
    it_mixed = spmv32_bytes(info["nnz"], N, NA) + 44 * N
    # bytes one product moves in the operator's own layout (stencil: slot
    # values, slot offsets, per-slice tail records and tail entries, the
    # input once, the output once): no column index per entry, so fewer
    # than the CSR count
    lay64, lay32 = A.product_bytes(), A.product_bytes(fp32=True)
    # the FP64 solve multiplies a bitwise symmetric stencil operator through
    # its symmetric form (the slots of offsets >= 0): its own bytes
    sym_state, sym64 = A.sym_state()
    own64 = sym64 if sym_state == 1 else lay64
    for leg, byts, lay in ((fp64, it_fp64, own64 + 88 * N), (mixed, it_mixed, lay32 + 44 * N)):
        ms_it = leg["ms"] / max(1, leg["iters"])
        leg["ms_per_iter"] = ms_it
        leg["iter_alg_bytes"] = byts  # CSR count of the product + the vector streams
        leg["iter_gbs"] = byts / (ms_it * 1e-3) / 1e9
        leg["iter_layout_bytes"] = lay  # the layout's own bytes + the vector streams
        leg["iter_layout_gbs"] = lay / (ms_it * 1e-3) / 1e9
    log(f"fp64 pcg {fp64['ms']:.1f} ms ({fp64['iters']} it, true {fp64['true_relres']:.2e}); "
        f"mixed {mixed['ms']:.1f} ms ({mixed['iters']} it, true {mixed['true_relres']:.2e}); "
        f"speedup {fp64['ms'] / mixed['ms']:.2f}x")
    A.destroy()
    dev.sync()

    ns = None
    ns_steps = 20 if args.ns_steps is None else args.ns_steps
    if not args.skip_ns and ns_steps > 0:
        ns_n = args.ns_n or n
        t0 = time.perf_counter()
        full = cavity.generate(ns_n, gamma=1e-3)
        log(f"generated full {ns_n}^3 cavity operators in {time.perf_counter() - t0:.1f} s")
        r64, f64 = ns_leg(dns, full, "fused", ns_steps, 1)
        r32, f32 = ns_leg(dns, full, "mixed", ns_steps, 1)
        # error of the mixed-precision run against the FP64 run, per field
        # (max abs difference over the inner cells / max abs FP64 value)
        N = full.n_inner
        err = {k: float(np.abs(f32[k][:N] - f64[k][:N]).max() / max(1e-300, np.abs(f64[k][:N]).max()))
               for k in f64}
        ns = {"n": ns_n, "re": 1000, "steps": ns_steps, "fp64": r64, "mixed": r32,
              "mixed_vs_fp64_rel_err": err}
        del f64, f32
        log(f"ns steps/s fp64 {ns['fp64']['steps_per_s']:.3f}, mixed {ns['mixed']['steps_per_s']:.3f}")
        del full

    cpu = None if args.skip_cpu else cpu_baseline(n, args.cpu_iters, args.seed, fp64["iters"])
    # the batched workload (64 uneven cavities), 10 timed steps, beside the
    # same instances stepped one after another
    batch = None if args.skip_batch else batch_workload(args, 11)

    # The dominant kernel of the path: the FP64 product inside the FP64 PCG
    # (its median launch, timed by events inside the solves), achieved = the
    # bytes the product's own layout moves per launch (over the symmetric
    # form: the values of the slots of offsets >= 0, slot offsets and codes,
    # tails, the input and output once).  The CSR count 12 nnz + 4 (rows+1)
    # + 8 rows + 8 cols per launch beside it is only a CSR equivalent: the
    # layout moves fewer bytes, so that rate can exceed the peak.
    us64, us32 = fp64["product_us"], mixed["product_us"]
    # operators of at most 2^19 rows form the direction inside the product
    if info["nrows"] <= 1 << 19:
        k64, k32, form = "k_iter_dirprod", "k_m_dirprod", "stencil layout"
    elif sym_state == 1:
        k64, k32, form = "k_iter_spmv_sym", "k_m_spmv", "symmetric form of the stencil layout"
    else:
        k64, k32, form = "k_iter_spmv<", "k_m_spmv", "stencil layout"
    ceiling = stream_ceiling()
    ach64 = own64 / (us64 * 1e-6) / 1e9
    traffic64 = measured_traffic(n, k64)
    roofline = {
        "bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(ach64 / HBM_PEAK_GBS, 4), "traffic": traffic64,
        "algorithmic_bytes": own64, "us_per_launch": round(us64, 2),
        "frac_traffic": round(traffic64 / (us64 * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if traffic64 else None,
        "csr_bytes": alg, "csr_equiv_gbs": round(alg / (us64 * 1e-6) / 1e9, 1), "full_layout_bytes": lay64,
        "kernel": f"{k64.rstrip('<')} (FP64 product of the FP64 PCG over the {form}, timed inside the solves), "
                  f"{n}^3"}
    # The dominant kernel of the timed mixed solve, its FP32 product: achieved
    # = the layout's bytes (the FP32 CSR count includes column indices this
    # layout never reads, and would put the rate above the peak).
    ach32 = lay32 / (us32 * 1e-6) / 1e9
    roofline_solve = {
        "bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(ach32 / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n, k32),
        "algorithmic_bytes": lay32, "us_per_launch": round(us32, 2), "csr_bytes": alg32,
        "kernel": f"{k32} (FP32 stencil-layout product of the mixed PCG, timed inside the solves), {n}^3"}

    out = {
        "metric": METRIC,
        "value": round(mixed["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": f"synthetic: generated {n}^3 uniform cavity Laplacian, seeded N(0,1) zero-mean RHS",
        "config": {"workload": f"{n}^3 cavity Poisson, Jacobi-PCG to 1e-10 FP64 relative true residual "
                               f"(mixed precision), one partition", "n": n, "nnz": info["nnz"],
                   "layout": hip.FORMAT_NAMES.get(info["format"], "?")},
        "pcg_fp64_ms": round(fp64["ms"], 3),
        "pcg_mixed_ms": round(mixed["ms"], 3),
        "mixed_speedup": round(fp64["ms"] / mixed["ms"], 3),
        # the speed-up splits into the cheaper FP32 iteration and the fewer
        # iterations of the deflated inner solve
        "mixed_speedup_per_iteration": round(fp64["ms_per_iter"] / mixed["ms_per_iter"], 3),
        "mixed_iteration_ratio": round(fp64["iters"] / max(1, mixed["iters"]), 3),
        "pcg_fp64": fp64,
        "pcg_mixed": mixed,
        # FP64: the CSR byte count of the metric, and the layout's own bytes;
        # FP32: the layout's own bytes (its CSR count includes column indices
        # the layout never reads, so that rate is only a CSR equivalent)
        "spmv_fp64": {"ms": spmv_ms, "gbs": spmv_gbs, "algorithmic_bytes": alg,
                      "layout_gbs": lay64 / (spmv_ms * 1e-3) / 1e9, "layout_bytes": lay64},
        "spmv_fp32": {"ms": spmv32_ms, "gbs": lay32 / (spmv32_ms * 1e-3) / 1e9, "layout_bytes": lay32,
                      "csr_equiv_gbs": alg32 / (spmv32_ms * 1e-3) / 1e9, "csr_bytes": alg32},
        "ns_steps": ns,
        "roofline": roofline,
        "roofline_solve": roofline_solve,
        "stream_ceiling_gbs": ceiling,
        "cpu_baseline": cpu,
        "batch": batch,
    }
    print(json.dumps(out), flush=True)


def summarize_pmc(fetch_dir, write_dir, kernels):
    """Per-dispatch HBM traffic of each named kernel from two rocprofv3 --pmc
    runs of `--probe spmv` (FETCH_SIZE, then WRITE_SIZE; counter CSVs under
    the two directories), scaled by what the same counters read for k_copy
    (2^27 doubles read and written per launch).  Returns
    {kernel: {"fetch_kib", "write_kib", "bytes"}} (medians over dispatches)
    and the two scale factors."""
    import csv

    def load(d, counter):
        rows = {}
        for f in pathlib.Path(d).rglob("*counter_collection.csv"):
            with open(f, newline="") as fh:
                for r in csv.DictReader(fh):
                    if r.get("Counter_Name") != counter:
                        continue
                    rows.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return rows

    fet, wri = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")

    def med(d, key):
        hits = [v for k, vs in d.items() if key in k for v in vs]
        return float(np.median(hits)) if hits else None

    copy_bytes = 8 * (1 << 27)
    f_scale = copy_bytes / 1024 / med(fet, "k_copy")
    w_scale = copy_bytes / 1024 / med(wri, "k_copy")
    out = {}
    for k in kernels:
        fk, wk = med(fet, k), med(wri, k)
        if fk is None or wk is None:
            continue
        out[k] = {"fetch_kib": fk, "write_kib": wk, "bytes": 1024 * (fk * f_scale + wk * w_scale)}
    return out, f_scale, w_scale


# kernel-name keys (substrings of the profiled names; "k_iter_dir(" is the
# four-launch form's direction sweep, not k_iter_dir_rows)
PCG_KERNELS = ("k_iter_dir_rows", "k_iter_dir(", "k_iter_spmv<", "k_iter_spmv_sym", "k_fold_two", "k_alpha_fin",
               "k_iter_update_rows",
               "k_iter_update_grid", "k_beta_fin", "k_m_dir", "k_m_spmv", "k_fold_dn", "k_m_update", "k_stencil<0")


def summary_traffic(args):
    """profiles/pcg_kernels_<n>_traffic.txt from the two --pmc runs of
    `--probe spmv --n <n>`."""
    n = args.n
    res, fs, ws = summarize_pmc(args.fetch, args.write, PCG_KERNELS)
    N = n ** 3
    lines = [
        f"# HBM traffic per launch of the {n}^3 PCG kernels (rocprofv3 PMC: TCC FETCH_SIZE and WRITE_SIZE in KiB per",
        "# dispatch, median over the dispatches of `bench.py --probe spmv`: 5 standalone FP64 products, 8",
        "# iterations of the FP64 PCG, 8 of the mixed PCG, 8 of the FP64 PCG without the symmetric form, 5 copies",
        "# of 2^27 doubles)",
        "#",
        "# commands (run from the repository root on one MI355X):",
        f"#   rocprofv3 --pmc FETCH_SIZE --output-format csv -d out/pmc_fetch -o run -- python3 -u bench.py "
        f"--probe spmv --n {n}",
        f"#   rocprofv3 --pmc WRITE_SIZE --output-format csv -d out/pmc_write -o run -- python3 -u bench.py "
        f"--probe spmv --n {n}",
        "#   python3 bench.py --summarize traffic --fetch out/pmc_fetch --write out/pmc_write "
        f"--n {n}",
        "#",
        "# calibration: k_copy of 2^27 doubles (1 GiB read + 1 GiB written per launch, known):",
        f"#   FETCH_SIZE scale {fs:.4f}, WRITE_SIZE scale {ws:.4f} (bytes = 1024 KiB x scale)",
        "#",
        "# kernel            fetch KiB     write KiB      bytes/launch   bytes/row",
    ]
    for k in PCG_KERNELS:
        if k not in res:
            continue
        r = res[k]
        lines.append(f"#   {k:14s} {r['fetch_kib']:12.1f} {r['write_kib']:12.1f} {r['bytes']:16.0f} "
                     f"{r['bytes'] / N:9.2f}")
    for k in PCG_KERNELS:
        if k in res:
            lines.append(f"traffic_bytes_per_launch[{k}]: {int(res[k]['bytes'])}")
    print("\n".join(lines))


def summary_kernels(args):
    """Per-kernel totals (calls, total/average/min/max duration, share) of a
    rocprofv3 --kernel-trace run, from its results database."""
    import sqlite3
    con = sqlite3.connect(args.db)
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows)
    lines = [f"# {line}" for line in args.title.splitlines()] + [
        "#",
        f"# summary: python3 bench.py --summarize kernels --db {args.db}",
        "# durations in microseconds (rocprofv3 kernel-trace timestamps)",
        "#",
        "# share   calls      total_us    avg_us    min_us    max_us  kernel"]
    for name, calls, tot, avg, mn, mx in rows:
        lines.append(f"{100 * tot / total:6.2f}% {calls:7d} {tot / 1e3:13.1f} {avg / 1e3:9.2f} {mn / 1e3:9.2f} "
                     f"{mx / 1e3:9.2f}  {name}")
    # roctx ranges (a --marker-trace run): host time per range message
    ranges = {}
    try:
        for ext, dur in con.execute("select extdata, duration from regions where category like 'MARKER%'"):
            msg = json.loads(ext or "{}").get("message", "?")
            ranges.setdefault(msg, []).append(dur / 1e3)
    except sqlite3.OperationalError:
        pass
    if ranges:
        lines += ["#", "# roctx ranges (host time, microseconds)", "#",
                  "# calls      total_us    avg_us    min_us    max_us  range"]
        for msg, d in ranges.items():
            lines.append(f"{len(d):7d} {sum(d):13.1f} {sum(d) / len(d):9.2f} {min(d):9.2f} {max(d):9.2f}  {msg}")
    print("\n".join(lines))


def summary_barrier(args):
    """profiles/device_barrier_cost.txt from a `--probe barrier` JSON line."""
    d = json.loads(pathlib.Path(args.json).read_text().strip().splitlines()[-1])
    lines = [
        "# Cost of one device-wide barrier on this MI355X (cooperative launch, every workgroup resident,",
        "# 2000 barriers per measurement, HIP events): relaxed device-scope atomic counter; with a device-scope",
        "# fence before arriving and after leaving (needed to publish data written before the barrier);",
        "# fenced with one double per workgroup stored before and all of them summed after.",
        "#",
        "# commands (run from the repository root on one MI355X):",
        "#   python3 bench.py --probe barrier > out/barrier.json",
        "#   python3 bench.py --summarize barrier --json out/barrier.json",
        "#",
        "# workgroups   relaxed_us   fenced_us   fenced_exchange_us"]
    for r in d["rows"]:
        lines.append(f"#   {r['workgroups']:6d} {r['relaxed_us']:12.3f} {r['fenced_us']:11.3f} {r['fenced_exchange_us']:14.3f}")
    print("\n".join(lines))


def summary_stream(args):
    """profiles/stream_ceiling.txt from a `--probe stream` JSON line."""
    d = json.loads(pathlib.Path(args.json).read_text().strip().splitlines()[-1])
    lines = [
        "# Practical HBM streaming ceiling of this MI355X: the best of each access mix over 2 GiB arrays",
        "# (8 GiB of HBM per test, far beyond the 256 MB Infinity Cache), HIP events over 10 launches.",
        "#",
        "# commands (run from the repository root on one MI355X):",
        "#   python3 bench.py --probe stream > out/stream.json",
        "#   python3 bench.py --summarize stream --json out/stream.json",
        "#",
        "# mix    bytes/access  nontemporal  grid   ms/launch   GB/s",
    ]
    for r in d["all"]:
        lines.append(f"#   {r['kind']:6s} {r['bytes_per_access']:4d} {str(r['nontemporal']):>12s} {r['grid']:6d} "
                     f"{r['ms']:10.4f} {r['gbs']:8.1f}")
    for k, r in d["best"].items():
        lines.append(f"best_gbs[{k}]: {r['gbs']}")
    print("\n".join(lines))


def probe_barrier(args):
    """Microseconds per device-wide barrier (gip_barrier_probe) for 8 to
    1024 resident workgroups: relaxed counter, with fences, with fences and
    one value per workgroup exchanged."""
    from gip_amd import hip
    dev = hip.Device(0)
    rows = []
    for grid in (8, 64, 128, 256, 512, 1024):
        r = {"workgroups": grid}
        for mode, name in ((0, "relaxed_us"), (1, "fenced_us"), (2, "fenced_exchange_us")):
            r[name] = round(hip.barrier_probe(dev, grid, 2000, mode), 3)
        rows.append(r)
        log(f"barrier, {grid:5d} workgroups: {r}")
    print(json.dumps({"probe": "barrier", "rows": rows}), flush=True)


def probe_stream(args):
    """GB/s of each streaming access mix over 2^28-double (2 GiB) arrays;
    bytes per launch: copy 16 n, read 8 n, write 8 n, mix 40 n."""
    from gip_amd import hip
    dev = hip.Device(0)
    n = 1 << 28
    bufs = (dev.array(n), dev.array(n), dev.array(n), dev.array(1 << 22))
    for a in bufs[:3]:
        hip.fill(a, 1.0)
    per = {"copy": 16, "read": 8, "write": 8, "mix": 40}
    rows = []
    for kind in ("copy", "read", "write", "mix"):
        for vec in (1, 2):
            for nt in (False, True):
                for grid in (0, 1024, 2048, 4096, 8192, 16384):
                    ms = hip.stream_probe(dev, kind, n, vec=vec, nontemporal=nt, grid=grid, reps=10, bufs=bufs)
                    gbs = per[kind] * n / (ms * 1e-3) / 1e9
                    rows.append({"kind": kind, "bytes_per_access": 8 * vec, "nontemporal": nt, "grid": grid,
                                 "ms": round(ms, 4), "gbs": round(gbs, 1)})
                    log(f"{kind:5s} {8 * vec:2d}B nt={int(nt)} grid={grid:5d}: {ms:.4f} ms {gbs:7.1f} GB/s")
    best = {k: max((r for r in rows if r["kind"] == k), key=lambda r: r["gbs"]) for k in per}
    print(json.dumps({"probe": "stream", "n_doubles": n, "best": best, "all": rows}), flush=True)


def probe_spmv(args):
    """Five FP64 SpMVs on the n^3 Laplacian, eight iterations of the FP64 and
    of the mixed PCG on it (their in-solve products), eight of the FP64 PCG
    without the symmetric form, and five copies of 2^27
    doubles (1 GiB read + 1 GiB written each): the launches a rocprofv3 --pmc
    run counts, the copy calibrating FETCH_SIZE/WRITE_SIZE against known
    bytes."""
    from gip_amd import cavity, hip
    cav = cavity.generate(args.n, gamma=1e-3, ops=cavity.OP_BITS["L"])
    L = cav.op("L")
    dev = hip.Device(0)
    A = hip.Matrix.from_csr(dev, L)
    info = A.info()
    x = dev.to_device(np.random.default_rng(1).standard_normal(cav.n_all))
    y = dev.array(cav.n_all)
    for _ in range(5):
        A.spmv(x, y)
    # 8 iterations of each PCG (their products: k_iter_spmv_sym over the
    # symmetric form, k_m_spmv), and 8 of the FP64 PCG on the operator
    # built without the symmetric form (k_iter_spmv)
    A32 = hip.Matrix.from_csr(dev, L, fp32=True)
    solver = hip.PCG(A32, cav.n_all)
    b = dev.to_device(np.concatenate([cavity.rhs_normal(cav.n_inner, 1), np.zeros(cav.n_all - cav.n_inner)]))
    z = dev.to_device(np.zeros(cav.n_all))
    for prec in (hip.PCG_FP64, hip.PCG_MIXED):
        solver.solve(b, y, z, maxit=8, tol=1e-10, precision=prec, poll=32)
    solver.destroy()
    A32.destroy()
    F = hip.Matrix.from_csr(dev, L, sym=False)
    solver = hip.PCG(F, cav.n_all)
    solver.solve(b, y, z, maxit=8, tol=1e-10, poll=32)
    solver.destroy()
    F.destroy()
    nc = 1 << 27
    a, b = dev.array(nc), dev.array(nc)
    hip.fill(a, 1.0)
    for _ in range(5):
        hip.copy(a, b)
    dev.sync()
    print(json.dumps({"probe": "spmv", "n": args.n, "spmv_algorithmic_bytes": spmv_bytes(info["nnz"], cav.n_inner,
                                                                                          cav.n_all),
                      "copy_bytes_read": 8 * nc, "copy_bytes_written": 8 * nc}), flush=True)


def batch_suite(args):
    """64 cavities of uneven size and Reynolds number, 50 steps: the batched
    integrator (one block-diagonal system, segmented kernels, hipGraph) vs
    the same instances stepped one after another with the single-cavity
    path.  Both solve the pressure to 1e-8 with the mixed-precision PCG."""
    print(json.dumps(batch_workload(args, 50, sequential=not args.skip_sequential, eager=args.batch_eager)),
          flush=True)


def batch_workload(args, steps, sequential=True, eager=False):
    """The 64-cavity batch for `steps` steps (the first untimed), and the
    same instances stepped one after another when `sequential`."""
    from gip_amd import batch as gbatch
    from gip_amd import cavity, dns, hip
    rng = np.random.default_rng(args.seed)
    B = 64
    sizes = [int(v) for v in rng.integers(32, 97, B)]
    re = np.exp(rng.uniform(np.log(100.0), np.log(3200.0), B))
    t0 = time.perf_counter()
    cavs = [cavity.generate(n, gamma=1.0 / r) for n, r in zip(sizes, re)]
    cells = sum(c.n_inner for c in cavs)
    log(f"generated {B} cavities ({cells} cells) in {time.perf_counter() - t0:.1f} s")
    dev = hip.Device(0)
    bt = gbatch.Batch(dev, cavs, 1.0 / re, pcg_tol=1e-8, pcg_maxit=20000, precision=hip.PCG_MIXED, poll=16,
                      use_graph=not eager, autotune=True)
    bt.step()  # warm-up: graph capture
    dev.sync()
    phases = dict.fromkeys(bt.PHASES, 0.0)
    t0 = time.perf_counter()
    for _ in range(steps - 1):
        bt.step()
        for k, v in bt.phase_ms().items():
            phases[k] += v
    dev.sync()
    t_batch = time.perf_counter() - t0
    log("batch phases, ms per step: " + ", ".join(f"{k} {v / (steps - 1):.2f}" for k, v in phases.items()))
    bt_u = [bt.field(i, "u") for i in range(B)]
    bt_info = [bt.info(i) for i in range(B)]
    bt.close()
    log(f"batch: {steps - 1} steps in {t_batch:.2f} s")
    t_seq, max_diff = float("nan"), float("nan")
    for i, (c, r) in enumerate(zip(cavs, re)):
        if not sequential:
            break
        if i == 0:
            t_seq, max_diff = 0.0, 0.0
        sim = dns.Simulation(cavity=c, gamma=1.0 / r, pcg_mode="mixed", pcg_tol=1e-8, pcg_maxit=20000)
        sim.step()
        sim.sync()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            sim.step()
        sim.sync()
        t_seq += time.perf_counter() - t0
        u = sim.field("u")
        ref = max(1e-30, float(np.abs(u).max()))
        max_diff = max(max_diff, float(np.abs(u - bt_u[i]).max()) / ref)
        sim.close()
    log(f"sequential: {t_seq:.2f} s; max relative u difference batch vs sequential {max_diff:.2e}")
    inst_steps = B * (steps - 1)
    dev.close()
    return {
        "metric": "batched cavities: instance-steps/s (64 uneven cavities, PCG to 1e-8, mixed precision)",
        "value": inst_steps / t_batch, "unit": "instance-steps/s", "n_gpus": 1, "steps": steps - 1, "warmup": 1,
        "higher_is_better": True,
        "sequential_instance_steps_per_s": inst_steps / t_seq if sequential else None,
        "speedup_vs_sequential": t_seq / t_batch if sequential else None,
        "max_rel_u_diff_vs_sequential": max_diff if sequential else None,
        "cells_total": cells, "sizes": sizes, "re": [round(float(v), 1) for v in re],
        "pcg_iters_last_step": [i["pcg_iters"] for i in bt_info],
        "graph_replay": not eager,
        "phase_ms_per_step": {k: v / (steps - 1) for k, v in phases.items()},
        "data": "synthetic: generated uniform cavities, lid u=1, from rest"}


def clustered_suite(args):
    """128^3 cavity on the tanh-clustered mesh x_i = (1 + tanh(beta (2i/n - 1))
    / tanh(beta)) / 2, beta = 1.5, Re = 1000: 100 Navier-Stokes steps from
    rest with the pressure solved to 1e-10 (FP64 and mixed precision), plus
    the L SpMV rate and both PCG solvers on a seeded zero-mean N(0,1)
    right-hand side (the 128^3 microbenchmarks)."""
    from gip_amd import cavity, dns, hip
    n, steps = 128, 100
    t0 = time.perf_counter()
    cav = cavity.generate(n, beta=1.5, gamma=1e-3)
    log(f"generated clustered {n}^3 cavity in {time.perf_counter() - t0:.1f} s")
    L = cav.op("L")
    N, NA = cav.n_inner, cav.n_all
    dev = hip.Device(0)
    A = hip.Matrix.from_csr(dev, L, fp32=True)
    info = A.info()
    x = dev.to_device(np.random.default_rng(1).standard_normal(NA))
    y = dev.array(NA)
    spmv_ms = time_spmv(hip, dev, A, x, y, 50)
    alg = spmv_bytes(info["nnz"], N, NA)
    b = np.zeros(NA)
    b[:N] = cavity.rhs_normal(N, args.seed)
    bnorm = float(np.linalg.norm(b[:N]))
    b_dev = dev.to_device(b)
    fp64 = pcg_leg(hip, dev, A, b_dev, N, NA, hip.PCG_FP64, 2, 1, 20000, bnorm)
    mixed = pcg_leg(hip, dev, A, b_dev, N, NA, hip.PCG_MIXED, 2, 1, 20000, bnorm)
    A.destroy()
    del x, y, b_dev
    r64, f64 = ns_leg(dns, cav, "fused", steps, 1)
    r32, f32 = ns_leg(dns, cav, "mixed", steps, 1)
    err = {k: float(np.abs(f32[k][:N] - f64[k][:N]).max() / max(1e-300, np.abs(f64[k][:N]).max())) for k in f64}
    log(f"clustered 128^3: spmv {alg / (spmv_ms * 1e-3) / 1e9:.0f} GB/s; pcg fp64 {fp64['ms']:.1f} ms "
        f"({fp64['iters']} it), mixed {mixed['ms']:.1f} ms ({mixed['iters']} it); ns steps/s fp64 "
        f"{r64['steps_per_s']:.2f}, mixed {r32['steps_per_s']:.2f}")
    print(json.dumps({
        "metric": "clustered 128^3 cavity (beta 1.5), Re=1000, PCG to 1e-10: Navier-Stokes steps/s, mixed precision",
        "value": r32["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "higher_is_better": True, "ns_fp64": r64, "ns_mixed": r32, "mixed_vs_fp64_rel_err": err,
        "spmv_fp64": {"ms": spmv_ms, "gbs": alg / (spmv_ms * 1e-3) / 1e9, "algorithmic_bytes": alg,
                      "layout": hip.FORMAT_NAMES.get(info["format"], "?")},
        "pcg_fp64": fp64, "pcg_mixed": mixed,
        "data": "synthetic: generated tanh-clustered 128^3 cavity, lid u=1, from rest; seeded N(0,1) zero-mean RHS"}),
        flush=True)


def reference_suite(args):
    """The reference's GIP_DNS::integrate over its Multicore specification
    (oracle/_ref/gip_ref_harness dns: the reference sources compiled -O2,
    one MPI rank, one core, timed with MPI_Wtime around integrate()) and the
    GPU integrator (the HIP specification through include/gip_dns.h, with
    the reference's PCG sequence and with the device-resident PCG; and the
    whole step device-resident and replayed from hipGraphs, as a
    one-instance batch), on the same one-partition cases with the
    reference's fixed 100 PCG iterations."""
    import tempfile
    from gip_amd import batch as gbatch
    from gip_amd import cavity, dns, hip
    exe = ROOT / "oracle" / "_ref" / "gip_ref_harness"
    if not exe.exists():
        log("reference harness absent (built where the reference sources are present)")
        return
    tmp = pathlib.Path(tempfile.mkdtemp(prefix="gip_ref_"))
    cases = []
    if (ROOT / "data" / "InputFiles" / "Topos" / "C100K" / "C100K_1p").exists():
        cases.append(("C100K", ROOT / "data", "C100K", 1e-3, 10))
    for n, g, steps in ((46, 1e-2, 10), (96, 1e-3, 3)):
        cavity.generate(n, gamma=g).write(tmp, f"cav{n}")
        cases.append((f"cavity {n}^3", tmp, f"cav{n}", g, steps))
    rows = []
    for label, root, name, gamma, steps in cases:
        out = subprocess.run([str(exe), "dns", name, str(steps), str(gamma), "100"], cwd=root, capture_output=True,
                             text=True, timeout=900, env=dict(os.environ, OMP_NUM_THREADS="1"))
        cpu_ms = None
        for line in out.stderr.splitlines():
            if line.startswith("integrate_ms"):
                cpu_ms = float(line.split()[1])
        res = {"case": label, "steps": steps, "cpu_ms_per_step": cpu_ms / steps if cpu_ms else None}
        u_fused = None
        for mode in ("reference", "fused"):
            sim = dns.Simulation(root=root, name=name, gamma=gamma, pcg_mode=mode)
            sim.step()  # first step: uploads and lazy set-up out of the timed region
            sim.sync()
            t0 = time.perf_counter()
            for _ in range(steps - 1):
                sim.step()
            sim.sync()
            res[f"gpu_{mode}_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / max(1, steps - 1)
            if mode == "fused":
                u_fused = sim.field("u")
            sim.close()
        dev = hip.Device(0)
        bt = gbatch.Batch(dev, [cavity.read(root, name)], [gamma], pcg_tol=1e-300, pcg_maxit=100,
                          precision=hip.PCG_FP64, poll=100, use_graph=True)
        bt.step()  # graph capture
        dev.sync()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            bt.step()
        dev.sync()
        res["gpu_graph_step_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / max(1, steps - 1)
        u_graph = bt.field(0, "u")
        res["graph_vs_fused_max_rel_u"] = float(np.abs(u_graph - u_fused).max() / max(1e-30, np.abs(u_fused).max()))
        bt.close()
        dev.close()
        if name == "C100K" and (root / "InputFiles" / "Topos" / "C100K" / "C100K_4p").exists():
            # the reference's own decomposition: four ranks, four cores
            mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
            if pathlib.Path(mpirun).exists():
                out4 = subprocess.run([mpirun, "-np", "4", str(exe), "dns", name, str(steps), str(gamma), "100"],
                                      cwd=root, capture_output=True, text=True, timeout=900,
                                      env=dict(os.environ, OMP_NUM_THREADS="1"))
                for line in out4.stderr.splitlines():
                    if line.startswith("integrate_ms"):
                        res["cpu_4ranks_ms_per_step"] = float(line.split()[1]) / steps
        log(f"{label}: cpu {res['cpu_ms_per_step']} ms/step (4 ranks: {res.get('cpu_4ranks_ms_per_step')}), "
            f"gpu reference-sequence "
            f"{res['gpu_reference_ms_per_step']:.2f}, device PCG {res['gpu_fused_ms_per_step']:.2f}, "
            f"graph step {res['gpu_graph_step_ms_per_step']:.2f} ms/step "
            f"(u vs device PCG {res['graph_vs_fused_max_rel_u']:.1e})")
        rows.append(res)
    print(json.dumps({
        "metric": "integrator ms per step, reference CPU path (1 core) vs GPU, fixed 100 PCG iterations",
        "value": rows[-1]["gpu_fused_ms_per_step"], "unit": "ms/step", "n_gpus": 1, "higher_is_better": False,
        "cases": rows, "cpu": {"cores": "1 (and 4 for the four-rank C100K row)", "kind": "reference",
                               "host": os.uname().nodename, "cpu_count": os.cpu_count()},
        "data": "C100K (the reference's shipped case, merged to one partition) and generated cavities"}), flush=True)


def cpu_suite(args):
    """The CPU-baseline plan: the reference's GIP_DNS::integrate over its
    Multicore specification (oracle/_ref/gip_ref_harness_O0: the as-shipped
    flags, no -O; _O3: -O3 -march=x86-64-v3), MPI_Wtime around integrate(),
    at 1 rank (1 core) and 4 ranks (4 cores, the reference's own
    decomposition), on the generated 46^3 cavity (Re=100; its 4-partition
    files from gip_cavity_split) and the shipped C100K (Re=1000; 4 partitions
    as shipped, 1 partition merged), 10 steps of fixed 100 PCG iterations;
    beside them the GPU: one partition (device PCG) and the same 4
    partitions on one device."""
    import tempfile
    from gip_amd import cavity, dns
    exes = {b: ROOT / "oracle" / "_ref" / f"gip_ref_harness_{b}" for b in ("O0", "O3")}
    mpirun = shutil.which("mpirun", path="/opt/conda/bin") or shutil.which("mpirun")
    if not all(e.exists() for e in exes.values()) or not mpirun:
        log("reference harness builds or MPI absent (built where the reference sources are present)")
        return
    tmp = pathlib.Path(tempfile.mkdtemp(prefix="gip_cpu_"))
    cav = cavity.generate(46, gamma=1e-2)
    cav.write(tmp, "cav46")
    cav.write_partitions(tmp, "cav46", 4)
    cases = [("cavity 46^3, Re=100", tmp, "cav46", 1e-2)]
    if (ROOT / "data" / "InputFiles" / "Topos" / "C100K" / "C100K_4p").exists():
        cases.append(("C100K, Re=1000", ROOT / "data", "C100K", 1e-3))
    steps, rows = 10, []
    for label, root, name, gamma in cases:
        res = {"case": label, "steps": steps}
        for build, exe in exes.items():
            for np_ in (1, 4):
                cmd = [str(exe), "dns", name, str(steps), str(gamma), "100"]
                if np_ > 1:
                    cmd = [mpirun, "-np", str(np_)] + cmd
                out = subprocess.run(cmd, cwd=root, capture_output=True, text=True, timeout=900,
                                     env=dict(os.environ, OMP_NUM_THREADS="1"))
                ms = None
                for line in out.stderr.splitlines():
                    if line.startswith("integrate_ms"):
                        ms = float(line.split()[1]) / steps
                res[f"cpu_{build}_np{np_}_ms_per_step"] = ms
        sim = dns.Simulation(root=root, name=name, gamma=gamma, pcg_mode="fused")
        sim.step()
        sim.sync()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            sim.step()
        sim.sync()
        res["gpu_1partition_device_pcg_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / (steps - 1)
        sim.close()
        parts = dns.PartitionedSimulation(root=root, name=name, nparts=4, gamma=gamma)
        parts.step()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            parts.step()
        res["gpu_4partitions_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / (steps - 1)
        parts.close()
        log(f"{label}: " + ", ".join(f"{k} {v:.2f}" for k, v in res.items() if isinstance(v, float)))
        rows.append(res)
    print(json.dumps({
        "metric": "integrator ms per step: the reference CPU path at 1 and 4 ranks, -O0 and -O3, beside the GPU",
        "value": rows[0]["gpu_1partition_device_pcg_ms_per_step"], "unit": "ms/step", "n_gpus": 1,
        "higher_is_better": False, "cases": rows,
        "cpu": {"kind": "reference", "cores": "1 per rank (1 and 4 ranks)", "host": host_cpu(),
                "builds": {"O0": "as shipped (compile_cpu.sh: no -O)", "O3": "-O3 -march=x86-64-v3"},
                "timer": "MPI_Wtime around integrate()"},
        "data": "generated 46^3 cavity (1 and 4 partitions) and the shipped C100K (4 partitions; 1 merged)"}),
        flush=True)


def ns256_suite(args):
    """The mixed-precision configuration's Navier-Stokes check: --ns-steps
    steps (default here 2) of the 256^3 cavity at Re=1000 with the pressure
    solved to 1e-10, on the GPU with the mixed-precision and the FP64
    device PCG, and on the FP64 oracle (the C restatement of the reference
    CPU path, its row loops on the host's cores and its dot products summed
    in 4096-term blocks); errors per field against the oracle: u, v, w
    against the norm of the whole velocity field and against their own, p
    mean-free over the inner cells (its constant mode is the pinned
    Laplacian's near-null mode, fixed only loosely by any solver stopped at
    1e-10; no velocity depends on it)."""
    from gip_amd import cavity, dns
    sys.path.insert(0, str(ROOT))
    from oracle import pyoracle
    n, steps = args.n, 2 if args.ns_steps is None else args.ns_steps
    t0 = time.perf_counter()
    cav = cavity.generate(n, gamma=1e-3)
    N = cav.n_inner
    log(f"generated {n}^3 cavity in {time.perf_counter() - t0:.1f} s")
    out = {}
    for mode in ("mixed", "fused"):
        sim = dns.Simulation(cavity=cav, gamma=1e-3, pcg_mode=mode, pcg_tol=1e-10, pcg_maxit=20000)
        rows = []
        for k in range(steps):
            g = sim.step()
            rows.append(g)
            log(f"gpu {mode} step {k + 1}: pcg {g['pcg_iters']} massr {g['massr']:.3e}")
        out[mode] = {"steps": rows, "fields": {f: sim.field(f) for f in ("u", "v", "w", "p")}}
        sim.close()
    pyoracle.set_threads(min(16, len(os.sched_getaffinity(0))))
    pyoracle.set_blocked_dot(True)
    orc = pyoracle.Dns(cavity=cav, gamma=1e-3, pcg_maxit=20000, pcg_tol=1e-10)
    orows = []
    for k in range(steps):
        t0 = time.perf_counter()
        w = orc.step()
        orows.append(w)
        log(f"oracle step {k + 1}: pcg {w['pcg_iters']} massr {w['massr']:.3e} ({time.perf_counter() - t0:.0f} s)")
    want = {f: orc.field(f) for f in ("u", "v", "w", "p")}
    orc.close()
    vel = float(np.sqrt(sum(np.linalg.norm(want[f]) ** 2 for f in ("u", "v", "w"))))
    wp = want["p"][:N] - want["p"][:N].mean()
    res = {}
    for mode in ("mixed", "fused"):
        got = out[mode]["fields"]
        gp = got["p"][:N] - got["p"][:N].mean()
        res[mode] = {
            "vs_velocity_norm": {f: float(np.linalg.norm(got[f] - want[f]) / vel) for f in ("u", "v", "w")},
            "own_norm": {f: float(np.linalg.norm(got[f] - want[f]) / np.linalg.norm(want[f])) for f in ("u", "v", "w")},
            "p_mean_free": float(np.linalg.norm(gp - wp) / np.linalg.norm(wp)),
            "time_rel": abs(out[mode]["steps"][-1]["time"] - orows[-1]["time"]) / orows[-1]["time"],
            "pcg_iters": [r["pcg_iters"] for r in out[mode]["steps"]]}
        log(f"{mode} vs FP64 oracle: {res[mode]}")
    print(json.dumps({
        "metric": f"{n}^3 cavity, Re=1000, pressure to 1e-10: fields after {steps} steps against the FP64 oracle",
        "value": res["mixed"]["p_mean_free"], "unit": "relative L2 (p, mean-free)", "n_gpus": 1,
        "higher_is_better": False, "steps": steps, "errors": res,
        "oracle_pcg_iters": [r["pcg_iters"] for r in orows],
        "oracle": "C restatement of the reference CPU path; row loops threaded (bitwise the serial ones), dot "
                  "products in 4096-term blocks",
        "data": f"synthetic: generated {n}^3 uniform cavity, lid u=1, from rest"}), flush=True)


def adjoint_suite(args):
    """64^3 cavity, Re=400, 200 steps with checkpoints every 20 steps:
    J = 1/2 sum V (u^2+v^2+w^2) at the end; gradient with respect to the
    64x64 lid values (1 + 0.05 N(0,1)) and the initial velocity, checked by
    central differences (eps 1e-6) along 16 random directions, dt frozen."""
    from gip_amd import adjoint as gadj
    from gip_amd import cavity, hip
    n, steps, every = 64, 200, 20
    cav = cavity.generate(n, gamma=1.0 / 400.0)
    dev = hip.Device(0)
    adj = gadj.Adjoint(dev, cav, steps=steps, checkpoint_every=every, gamma=1.0 / 400.0, pcg_tol=1e-13)
    rng = np.random.default_rng(args.seed)
    lid = 1.0 + 0.05 * rng.standard_normal(n * n)
    u0, v0, w0 = (np.zeros(cav.n_inner) for _ in range(3))
    adj.forward(lid, u0, v0, w0)  # freezes dt
    t0 = time.perf_counter()
    adj.forward(lid, u0, v0, w0)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    adj.backward()
    t_bwd = time.perf_counter() - t0
    J, ad, fd, rel = gadj.fd_check(adj, lid, u0, v0, w0, ndir=16, eps=1e-6, seed=args.seed)
    adj.close()
    log(f"adjoint: forward {t_fwd:.2f} s, backward {t_bwd:.2f} s, max FD rel err {rel.max():.2e}")
    print(json.dumps({
        "metric": "adjoint of 200 cavity steps (64^3, Re=400): backward/forward time, FD agreement",
        "value": t_bwd / t_fwd, "unit": "backward/forward", "n_gpus": 1, "higher_is_better": False,
        "forward_s": t_fwd, "backward_s": t_bwd, "J": J, "fd_max_rel_err": float(rel.max()),
        "fd_rel_err": [float(v) for v in rel], "directional_adjoint": [float(v) for v in ad],
        "directional_fd": [float(v) for v in fd], "checkpoint_every": every, "steps": steps,
        "data": "synthetic: generated 64^3 uniform cavity, lid 1 + 0.05 N(0,1), from rest"}), flush=True)


if __name__ == "__main__":
    main()
