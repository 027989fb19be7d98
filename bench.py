#!/usr/bin/env python3
"""Seconds per gauge configuration of the sLapH contraction path on one MI355X.

One step = VdaggerV for every momentum and timeslice + all pipi I=2 diagrams
(C2c, C4cD, C4cB, C4cC over |P|^2 <= 4 and C6cD at P = 0) over all dilution
block pairs, on synthetic inputs already resident in HBM:
L = 32, T = 64, N_ev = 220, 27 momenta |p|^2 <= 3, 5 random vectors,
dilution TB16 / EI4 / DF4.  Other workloads (--workload): c2 (24^3 x 48,
N_ev = 120, D = 32), disp48 (48^3 x 96, N_ev = 400, displaced operators, one
eigenvector timeslice resident) and small (4^3 x 8).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pipi32|c2|disp48|small]

Prints one JSON line.  `roofline` is the largest kernel of the step, the
queued left-stationary trace kernel k_trace_lstat_q (C4cC traces), timed with
HIP events around each of its launches on the stream it runs on; its
algorithmic bytes per launch are the distinct product matrices the launch
reads plus the partial sums it writes.  `roofline_vdaggerv` is the VdaggerV
kernel k_vdv_ft on its own operation count against the FP64 peak (the
roofline object itself for workloads without C4cC).  cpu_baseline is the
C++ / OpenMP restatement of the reference's algorithm (oracle/cpu_ref.cpp)
timed on a bounded sample on this host and extrapolated (see cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "slaph-contractions_amd"))
sys.path.insert(0, ROOT)

PIPI = ("C2c", "C4cD", "C4cB", "C4cC", "C6cD")
WORKLOADS = {
    # headline, the pipi I=2 workload: 32^3 x 64, N_ev = 220, 27 momenta
    # |p|^2 <= 3, TB16 / EI4 / DF4, C2c C4cD C4cB C4cC (|P|^2 <= 4) + C6cD (P = 0)
    "pipi32": dict(L=32, T=64, N=220, nrnd=5, dilT=16, dilE=4, pmax="0,1,2,3", diagrams=PIPI),
    # plain FP64 at production size: 24^3 x 48, N_ev = 120, 33 momenta
    # |p|^2 <= 4, TB4 / EI8 / DF4 (D = 32), C2c C3c C4cD C4cB
    "c2": dict(L=24, T=48, N=120, nrnd=5, dilT=4, dilE=8, pmax="0,1,2,3,4", diagrams=("C2c", "C3c", "C4cD", "C4cB")),
    # large volume with displaced operators: 48^3 x 96, N_ev = 400, 33 momenta
    # |p|^2 <= 4 undisplaced and one-link displaced in x, y, z, TB24 / EI4 /
    # DF4, C2c (both operator lists) and C4cD; one eigenvector timeslice
    # resident, seeded SU(3) links smeared by 3 stout steps at rho = 0.1
    "disp48": dict(L=48, T=96, N=400, nrnd=4, dilT=24, dilE=4, pmax="0,1,2,3,4", diagrams=("C2c", "C4cD"),
                   displaced=">x,>y,>z", streaming=True),
    # tiny CPU-checkable case
    "small": dict(L=4, T=8, N=8, nrnd=4, dilT=2, dilE=4, pmax="0,1", diagrams=PIPI),
}

PEAK_HBM = 8000.0      # GB/s
PEAK_FP64 = 78.6       # TFLOP/s, vector and MFMA

_CORR = {
    "C2c": "C2c:Q0:Op0:Q0:Op0:P0,1,2,3,4",
    "C3c": "C3c:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4",
    "C4cD": "C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4",
    "C4cB": "C4cB:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4",
    "C4cC": "C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4",
    "C6cD": "C6cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0",
}


def infile(w, pmax=None, diagrams=None, total_momenta=None):
    """Input file of a workload; pmax / diagrams / total_momenta override the
    operator momenta, the diagram list and the total momenta of every list."""
    lines = ["correlator_list = " + _CORR[d] for d in (diagrams or w["diagrams"])]
    ops = "operator_list = g5.d0.p" + (pmax or w["pmax"])
    if w.get("displaced"):  # second operator list: displaced operators, two-point functions
        ops += "\noperator_list = g5.d%s.p%s" % (w["displaced"], pmax or w["pmax"])
        lines.append("correlator_list = C2c:Q0:Op1:Q0:Op1:P0,1,2,3,4")
    if total_momenta is not None:
        lines = [ln.rsplit(":P", 1)[0] + ":P" + total_momenta for ln in lines]
    return """
Lt = {T}
Lx = {L}
Ly = {L}
Lz = {L}
number_of_eigen_vec = {N}
[quarks]
quark = u:{nrnd}:TB:{dilT}:EI:{dilE}:DF:4:synthetic
[operator_lists]
{ops}
[correlator_lists]
{corr}
""".format(ops=ops, corr="\n".join(lines), **w)


def trace_kernel(tuning=""):
    """The kernel of the roofline object: the crossed (C4cC) traces as
    symmetric products, or on the left-stationary kernel when SLAPH_SYM_OFF
    or the tuning (crossed_lstat=1) selects it (its block form under
    SLAPH_LSTAT_BLOCKS=1)."""
    if os.environ.get("SLAPH_SYM_OFF") or "crossed_lstat=1" in tuning.split():
        return "k_trace_lstat" if os.environ.get("SLAPH_LSTAT_BLOCKS") else "k_trace_lstat_q"
    return "k_trace_sym"


KERNEL_LABEL = {"k_trace_sym": "crossed traces as symmetric products",
                "k_trace_lstat_q": "non-separable traces", "k_trace_lstat": "non-separable traces"}


def kernel_source_hash(kernel):
    """sha256 (12 hex digits) of the source text of `kernel`'s definition in
    kernels.hip (from its __global__ line to the closing brace at column 0):
    a traffic figure measured for another version of the kernel is stale."""
    import hashlib
    src = open(os.path.join(ROOT, "slaph-contractions_amd", "csrc", "kernels.hip")).read().splitlines()
    start = next((i for i, ln in enumerate(src) if "__global__" in ln and (" %s(" % kernel) in ln + src[min(i + 1, len(src) - 1)]),
                 None)
    if start is None:
        return None
    end = next(i for i in range(start, len(src)) if src[i] == "}")
    return hashlib.sha256("\n".join(src[start:end + 1]).encode()).hexdigest()[:12]


def kernel_traffic(kernel, workload):
    """HBM bytes of one launch of `kernel`, measured with two rocprofv3 --pmc
    runs (FETCH_SIZE, WRITE_SIZE, copy-calibrated) and kept under profiles/,
    with the kernel source hash it was measured on: (bytes or None, note)."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None, "no measurement for this workload"
    vals = {}
    for line in open(path):
        if not line.startswith("#") and line.strip():
            k, v = line.split(None, 1)
            vals[k] = v.strip()
    now = kernel_source_hash(kernel)
    if vals.get("kernel_sha") != now:
        return None, "stale: measured on kernel source %s, now %s" % (vals.get("kernel_sha"), now)
    return float(vals["bytes_per_launch"]), "profiles/%s_%s_traffic.txt (kernel source %s)" % (kernel, workload, now)


def measure_traffic(kernel, workload, copy_kib):
    """HBM bytes per launch of `kernel` in this run's configuration, from two
    child runs of this bench under rocprofv3 --pmc (FETCH_SIZE, then
    WRITE_SIZE), each scaled by what the counter reads for the device-to-device
    copy of one eigenvector timeslice (copy_kib KiB) in the same run."""
    import csv
    import subprocess
    import tempfile
    out = tempfile.mkdtemp(prefix="slaph_traffic_", dir=os.environ.get("TMPDIR", "/tmp"))
    child = [sys.executable, os.path.abspath(__file__), "--workload", workload, "--steps", "1", "--warmup", "0",
             "--no-cpu-baseline", "--no-precision-forms", "--no-autotune", "--no-graphs"]
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", counter, "--"] + child
        print("# traffic: %s" % " ".join(cmd), file=sys.stderr, flush=True)
        subprocess.run(cmd, check=True, timeout=600, stdout=subprocess.DEVNULL)
        kern, copy = [], []
        for r in csv.DictReader(open(os.path.join(out, counter + "_counter_collection.csv"))):
            v = float(r["Counter_Value"])
            if kernel + "<" in r["Kernel_Name"] or kernel + "(" in r["Kernel_Name"]:
                kern.append(v)
            elif "copyBuffer" in r["Kernel_Name"] and abs(v - copy_kib) < 0.75 * copy_kib:
                copy.append(v)
        if not kern or not copy:
            raise RuntimeError("traffic: no %s dispatches of %s or of the calibration copy" % (counter, kernel))
        res[counter] = (sum(kern) / len(kern), copy_kib / (sum(copy) / len(copy)), len(kern))
    (f, sf, n), (w, sw, _) = res["FETCH_SIZE"], res["WRITE_SIZE"]
    return {"bytes_per_launch": (sf * f + sw * w) * 1024.0, "fetch_kib": f, "write_kib": w, "fetch_scale": sf,
            "write_scale": sw, "launches": n, "kernel": kernel, "kernel_sha": kernel_source_hash(kernel)}


def cpu_baseline(w, text, sample_slice_pairs=2):
    """The reference's algorithm as the C++ / OpenMP restatement
    (oracle/cpu_ref.cpp, checked against the reference's golden files) on this
    host, on a bounded sample of the same workload, extrapolated to one
    configuration:

    * contraction: every dilution block pair runs on its own thread (the
      reference's `omp for schedule(dynamic)` over block pairs,
      Correlators.cpp:69-116) for its first `sample_slice_pairs` slice pairs,
      so the factor builds of the keys they need and their traces are timed
      under the real concurrency; per block pair the time is then (factor keys
      it needs, by type) x (measured seconds per key of that type) + (its slice
      pairs) x (measured trace seconds per slice pair), and the block pairs are
      dealt to the threads in order, each to the first free thread;
    * VdaggerV: one (p, t) product timed on a K-slab of the lattice and scaled
      by K, times the momenta (and displacement paths) per timeslice, with the
      timeslices dealt to the threads (OperatorFactory.cpp:263).
    Inputs are random of the workload's shapes (values do not change the
    work)."""
    import numpy as np
    from oracle import cpu_ref
    from oracle import slaph_oracle as o
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    rng = np.random.default_rng(7)
    cfg = o.parse_infile(text)
    lut = o.init_lookup_tables(cfg)
    L, T, N, dilT = w["L"], w["T"], w["N"], w["dilT"]
    K = 3 * L ** 3
    nb = T // dilT
    cols = nb * w["dilE"] * 4
    perams = [rng.standard_normal((4 * N * T, cols)) + 1j * rng.standard_normal((4 * N * T, cols))
              for _ in range(w["nrnd"])]
    rnds = [(rng.integers(0, 2, T * N * 4) * 2 - 1 + 1j * (rng.integers(0, 2, T * N * 4) * 2 - 1)) / np.sqrt(2)
            for _ in range(w["nrnd"])]
    M = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    ref = cpu_ref.CpuRef(cfg, lut, lambda i, t: M, perams, rnds)
    bps = o.block_pairs(T, dilT)
    used = min(threads, len(bps))
    _, wall = ref.contract(threads=used, first_block_pair=0, block_pairs=len(bps), slice_pairs=sample_slice_pairs)
    st = ref.last_stats()
    per_key = {q: st[q][0] / st[q][1] for q in ("Q0", "Q1", "Q2") if st[q][1] > 0}
    n_sampled = sum(min(sample_slice_pairs, len(o.slice_pairs(a, b, T, nb))) for a, b in bps)
    factor_s = sum(st[q][0] for q in ("Q0", "Q1", "Q2"))
    per_pair = (st["busy"] - factor_s) / n_sampled  # thread-seconds of traces and resolves per slice pair
    times = []
    for a, b in bps:
        sp = o.slice_pairs(a, b, T, nb)
        keys = cpu_ref.factor_keys(lut, dilT, sp)
        tb = len(sp) * per_pair
        for q in ("Q0", "Q1", "Q2"):
            nq = sum(1 for k in keys if k[0] == q)
            if nq:
                tb += nq * per_key[q]
        times.append(tb)
    free = [0.0] * used
    for tb in times:  # schedule(dynamic): the next block pair to the first free thread
        i = free.index(min(free))
        free[i] += tb
    contract_s = max(free)
    # VdaggerV: one product on a K-slab, scaled by K
    slab = max(3, (K // 8) // 3 * 3)
    V = rng.standard_normal((slab, N)) + 1j * rng.standard_normal((slab, N))
    ph = np.exp(-2j * np.pi * rng.random(slab // 3))
    _, t_slab = cpu_ref.vdaggerv(V, ph)
    t_prod = t_slab * K / slab
    n_ops = sum(1 for vid, _, _ in lut["vdv_lookup"] if vid != lut["index_of_unity"])
    free = [0.0] * min(threads, T)
    for _ in range(T):
        i = free.index(min(free))
        free[i] += n_ops * t_prod
    vdv_s = max(free)
    total = vdv_s + contract_s
    sample = ("C++/OpenMP restatement of the reference (oracle/cpu_ref.cpp), random inputs of the workload's shapes: "
              "the first %d slice pairs of each of the %d block pairs, one block pair per thread on %d threads "
              "(%.1f s wall; factor keys %s s each; %.3f thread-s of traces per slice pair), extrapolated per "
              "block pair by its keys and slice pairs and scheduled over %d threads -> contraction %.0f s; "
              "VdaggerV: one product on %d of %d rows (%.2f s), x %d operators x %d timeslices over %d threads "
              "-> %.0f s" % (sample_slice_pairs, len(bps), used, wall,
                             json.dumps({q: round(v, 3) for q, v in per_key.items()}), per_pair, used, contract_s,
                             slab, K, t_slab, n_ops, T, min(threads, T), vdv_s))
    return {"value": total, "unit": "s", "cores": used, "kind": "port", "sample": sample,
            "cpu": _cpu_model(), "threads_available": threads}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _describe(name, w):
    corr = {"pipi32": "C2c C4cD C4cB C4cC (P^2<=4) + C6cD (P=0)", "small": "C2c C4cD C4cB C4cC (P^2<=4) + C6cD (P=0)",
            "c2": "C2c C3c C4cD C4cB (P^2<=4)",
            "disp48": "C2c (undisplaced and displaced >x,>y,>z) + C4cD (P^2<=4), one eigenvector timeslice "
                      "resident, SU(3) links + 3 stout steps rho=0.1 per config"}[name]
    return "%s: L=%d T=%d Nev=%d, %d rnd vecs, TB%d/EI%d/DF4, g5 p^2<=%s, %s, plain FP64 path" % (
        name, w["L"], w["T"], w["N"], w["nrnd"], w["dilT"], w["dilE"], w["pmax"][-1], corr)


def _kstat(ctx, kernel):
    """(ms, launches, bytes, flops) of `kernel` in the last step, or None."""
    try:
        return tuple(ctx.kernel_stat(kernel, k) for k in ("ms", "launches", "bytes", "flops"))
    except Exception:  # the kernel did not run in this workload
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="pipi32", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision-forms", action="store_true",
                    help="skip timing the precision forms against the FP64 path")
    ap.add_argument("--chunk-gib", type=float, default=0.0, help="chunk memory budget (0: automatic)")
    ap.add_argument("--no-graphs", action="store_true", help="launch the contraction directly (no HIP graph replay)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default schedules (otherwise the autotuner picks them before the timed steps)")
    ap.add_argument("--precision", default="fp64,fp64",
                    help="VdaggerV and factor forms of the timed run (reported in config; the FP64 path by default)")
    ap.add_argument("--measure-traffic", action="store_true",
                    help="measure roofline.traffic in this run: two child runs under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path "
              "(block pairs of slice pairs would be split across GPUs)", file=sys.stderr)
        return 2
    w = WORKLOADS[args.workload]
    traffic_run = None
    if args.measure_traffic:  # before this process touches the GPU
        traffic_run = measure_traffic(trace_kernel(), args.workload, 3 * w["L"] ** 3 * w["N"] * 16 / 1024.0)
    import slaph
    text = infile(w)
    ctx = slaph.Context(0)
    if w.get("streaming"):
        ctx.set_timeslice_streaming(True)  # before configure: V of one timeslice is allocated
    ctx.configure(text)
    ctx.set_chunk_budget(int(args.chunk_gib * (1 << 30)))
    ctx.set_graphs(not args.no_graphs)
    pv, pf = args.precision.split(",")
    ctx.set_precision(pv, pf)
    t0 = time.perf_counter()
    # streamed eigenvectors are generated per timeslice inside each step (in
    # place of the reference's per-timeslice file read), without the
    # orthonormalisation pass
    ctx.synthetic_inputs(20261203, not w.get("streaming"))
    print("# inputs generated in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    ops = slaph.OperatorFactory(ctx)

    def step():
        if w.get("displaced"):
            ctx.synthetic_gauge(20261204, 0.2)
            ctx.smear_stout(0.1, 3)
        ops.create_operators()
        slaph.contract(ctx)

    for i in range(args.warmup):
        step()
        print("# warmup %d done" % i, file=sys.stderr, flush=True)
    tune = None
    if not args.no_autotune:
        # default schedules first (the "before"), then the autotuner's choice
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
        before = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        report = ctx.autotune(2)
        tune = {"default_s_per_config": before, "chosen": report["chosen"],
                "tuning_s": time.perf_counter() - t0, "knobs": report["knobs"]}
        print("# autotuned in %.1f s: %s" % (tune["tuning_s"], report["chosen"]), file=sys.stderr, flush=True)
        step()
    stage = {"vdaggerv": 0.0, "factors": 0.0, "contract": 0.0}
    tkern = trace_kernel(ctx.tuning())
    stats = {tkern: [0.0] * 4, "k_vdv_ft": [0.0] * 4}
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        for k in stage:
            stage[k] += ctx.stage_time(k) / args.steps
        for kern in stats:
            st = _kstat(ctx, kern)
            if st is None:
                stats[kern] = None
            elif stats[kern] is not None:
                stats[kern] = [a + b for a, b in zip(stats[kern], st)]
        print("# step %d done" % i, file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1000.0
    L, T, N = w["L"], w["T"], w["N"]
    headline = args.workload == "pipi32"
    line = {
        "metric": ("s per gauge config (VdV+all diagrams) L=32 T=64 Nev=220: mixed vs FP64 @1e-10" if headline else
                   "s per gauge config (VdV+all diagrams) L=%d T=%d Nev=%d" % (L, T, N)),
        "value": ms / 1000.0,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64,fp64" else "mixed (%s)" % args.precision,
        "data": "synthetic (seeded, device-generated; %s)" % (
            "eigenvectors generated per timeslice, not orthonormalised" if w.get("streaming")
            else "orthonormalised eigenvectors"),
        "config": {"workload": _describe(args.workload, w), "precision": args.precision,
                   "hip_graphs": not args.no_graphs, "tuning": ctx.tuning()},
        "stages_ms": {"vdaggerv": stage["vdaggerv"], "factors": stage["factors"],
                      "contract_incl_factors": stage["contract"]},
    }
    ls = stats[tkern]
    if ls is not None:
        launches = max(ls[1], 1.0)
        kern_ms, bytes_per_launch = ls[0] / launches, ls[2] / launches
        achieved = bytes_per_launch / (kern_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0
        if traffic_run is not None and traffic_run.get("kernel", tkern) == tkern:
            traffic, tnote = traffic_run["bytes_per_launch"], "measured in this run: %s" % json.dumps(traffic_run)
        elif traffic_run is not None:
            traffic, tnote = None, "measured for %s, but the tuning ran %s" % (traffic_run["kernel"], tkern)
        else:
            traffic, tnote = kernel_traffic(tkern, args.workload)
        line["roofline"] = {"bound": "hbm", "kernel": "%s (%s)" % (tkern, KERNEL_LABEL[tkern]), "achieved": achieved,
                            "peak": PEAK_HBM, "unit": "GB/s", "frac": achieved / PEAK_HBM,
                            "traffic": traffic, "traffic_source": tnote,
                            "traffic_unit": "bytes per launch (HBM, rocprofv3 --pmc)",
                            "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_ms": kern_ms,
                            "launches_per_step": ls[1] / args.steps}
    vf = stats["k_vdv_ft"]
    if vf is not None:
        launches = max(vf[1], 1.0)
        kern_ms, flops_per_launch = vf[0] / launches, vf[3] / launches
        tf = flops_per_launch / (kern_ms / 1e3) / 1e12 if kern_ms > 0 else 0.0
        n_vdv = ctx.query("n_vdaggerv") - (1 if ctx.query("index_of_unity") >= 0 else 0)
        obj = {"bound": "fp64", "kernel": "k_vdv_ft (VdaggerV, sitewise Fourier form)", "achieved": tf,
               "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": tf / PEAK_FP64,
               "flops_per_launch": flops_per_launch, "avg_launch_ms": kern_ms,
               "launches_per_step": vf[1] / args.steps,
               "direct_product_equivalent_tflops": 8.0 * N * N * 3 * L ** 3 * n_vdv * T / (vf[0] / args.steps / 1e3)
               / 1e12 if vf[0] > 0 else None}
        if "roofline" in line:
            line["roofline_vdaggerv"] = obj
        else:
            line["roofline"] = obj
    if tune is not None:
        line["autotune"] = tune
    if not args.no_precision_forms and args.precision == "fp64,fp64":
        pf = precision_forms(ctx, ops, ms / 1000.0, step)
        line["precision_forms"] = pf
        ok = [f for f in pf if f["within_1e-10"]]
        # `value` is the plain FP64 path; a mixed form would replace it only if
        # every correlator stayed within 1e-10 of FP64
        line["fastest_mixed_form_within_1e-10"] = min(ok, key=lambda f: f["s_per_config"]) if ok else None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, text)
    print(json.dumps(line))
    return 0


def precision_forms(ctx, ops, fp64_s, step):
    """Time each precision form (one warm-up and one timed configuration on the
    inputs of the FP64 run) and state its per-correlator relative error against
    the FP64 path: err = max_t |C(t) - C_fp64(t)| / max_t |C_fp64(t)|."""
    import numpy as np
    import slaph

    def correlators():
        return {d: ctx.correlators(d) for d in ctx.diagrams()}

    ref = correlators()
    out = []
    for vdv, fac in (("fp32", "fp64"), ("bf16x3", "fp64"), ("ozaki", "fp64"), ("fp64", "fp32"), ("fp64", "fp48"),
                     ("ozaki", "fp48"), ("fp32", "fp32")):
        ctx.set_precision(vdv, fac)
        step()
        t0 = time.perf_counter()
        step()
        s = time.perf_counter() - t0
        got = correlators()
        err = np.array([np.max(np.abs(got[d][n] - v)) / max(np.max(np.abs(v)), 1e-300)
                        for d, vals in ref.items() for n, v in vals.items()])
        out.append({"vdaggerv": vdv, "factors": fac, "s_per_config": s, "speedup_vs_fp64": fp64_s / s,
                    "max_rel_err": float(err.max()), "median_rel_err": float(np.median(err)),
                    "n_correlators": int(err.size), "within_1e-10": bool(err.max() <= 1e-10)})
        print("# precision form vdaggerv=%s factors=%s: %.3f s, max rel err %.2e" % (vdv, fac, s, err.max()),
              file=sys.stderr, flush=True)
    ctx.set_precision("fp64", "fp64")
    return out


if __name__ == "__main__":
    sys.exit(main())
