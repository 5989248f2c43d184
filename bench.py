"""FRI-CCSD iterations/s on one MI355X (GPU) at o=32, v=256, m=4e6.

One step = one FRI-CCSD iteration of the reference's kernel loop
(fricc/friccsd.py:416-430): update_amps (two FRI compressions of T2, the 11
sparse contractions plus the ladder's t1 dressing, the dense remainder,
denominators), the amplitude-change norm and the energy.  Workload: the
synthetic density-fitted system (nocc=32, nvir=256, naux=864, seed 1;
fricc.systems), MP2 guess, systematic FRI with m = 4e6 uniforms from
srand(1)/rand(), level shift 2.0 (plain Jacobi diverges on this system), all
state resident in HBM before the timed region.

The headline value is the "mixed" precision mode (every sparse contraction
on BF16 MFMA with hi/lo split operands -- three passes, FP32 accumulate --
the dense products in FP32; compression, energy and denominators in FP64),
over c3's 40 iterations: W warmup iterations, then K timed ones (bracketed by
synchronisation), then the rest of the 40, every iteration's latency
reported.  Timed beside it: c3's mixture as written ("bf16-ladder": the
ladder on single-pass BF16, everything else FP32), FP32 and the FP64 plain
form, each with the per-term residual error against the FP64 step on the
same sampled T2 (the single-pass ladder misses the |dE| <= 1e-6 Ha target;
the headline mode meets it).  Prints one JSON line.

``--summarize-profile DIR`` (no GPU) turns a rocprofv3 --kernel-trace
--stats output directory into the markdown tables kept under profiles/.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fri-cc_amd"))

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_GBS = 8000.0
# precision mode -> (sparse contractions, ladder); mirrors fricc.friccsd.PRECISION_MODES
MODE_PREC = {"fp64": ("fp64", "fp64"), "fp32": ("fp32", "fp32"), "bf16-ladder": ("fp32", "bf16x1"),
             "mixed": ("bf16", "bf16")}


def _pct(a, q):
    import numpy as np

    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def run_mode(eris, cfg, mode, steps, warmup, seed, solver=None, total=None):
    """``total`` (>= warmup + steps) iterations from the MP2 guess: warmup
    iterations, the timed window of ``steps`` (bracketed by synchronisation:
    ``it_per_s``), then the rest; every iteration's host-clock latency (each
    includes the step's one read of energy and norm) and energy recorded, and
    per-section HIP-event times averaged over the window."""
    import torch

    from fricc.friccsd import FRICCSD, SectionTimer
    from fricc.uniforms import GlibcUniforms

    total = max(total or 0, warmup + steps)
    settings = {"m_keep": cfg["m_keep"], "compression": "fri", "sampling_method": "systematic",
                "precision": mode}
    cc = (solver or FRICCSD)(eris, settings, uniforms=GlibcUniforms(seed))
    cc.level_shift = 2.0
    t1, t2 = cc.get_init_guess(eris)
    lat, energies, window = [], [], None
    for it in range(total):
        if it == warmup:
            torch.cuda.synchronize()
            cc.timer = SectionTimer()
            t_win = time.perf_counter()
        ts = time.perf_counter()
        cc._iter = it
        t1, t2 = cc.update_amps(t1, t2, eris)  # energy and norm read back each cycle, as the loop does
        lat.append(1e3 * (time.perf_counter() - ts))
        energies.append(cc._last_step[2])
        if it == warmup + steps - 1:
            torch.cuda.synchronize()
            window = time.perf_counter() - t_win
            sec = cc.timer.ms()
            cc.timer = None
    info = cc.sparse_info[0]
    wl = lat[warmup:warmup + steps]
    return {"state": (t1, t2), "slots": cc._slots, "it_per_s": steps / window, "ms_per_step": 1e3 * window / steps,
            "p50_ms": _pct(wl, 50), "p99_ms": _pct(wl, 99),
            "sections_ms": {k: sum(v) / len(v) for k, v in sec.items() if v},
            "latency_ms": lat, "energies": energies, "n_D": info["n_D"], "n_sampled": info["n_sampled"]}


def per_term_errors(eris, cfg, mode, t1, t2):
    """One step from the same (t1, t2) with the same uniforms in FP64 and in
    ``mode``: both compress the same T2 (compression is FP64 in every mode),
    so the differences are the arithmetic alone.  Each compressed contraction
    is also evaluated into its own zeroed FP64 buffer (friccsd.TERM_PROBE);
    reported per term: max |d| / max |ref| and ||d||_F / ||ref||_F; then the
    same for t2new, t1new and the correlation energy."""
    import torch

    from fricc import friccsd
    from fricc.friccsd import FRICCSD
    from fricc.uniforms import FixedUniforms

    us = [0.3141592653589793, 0.2718281828459045]
    settings = {"m_keep": cfg["m_keep"], "compression": "fri", "sampling_method": "systematic"}
    probes, outs = {}, {}
    for m in ("fp64", mode):
        cc = FRICCSD(eris, dict(settings, precision=m), uniforms=FixedUniforms(us))
        cc.level_shift = 2.0
        friccsd.TERM_PROBE = probes.setdefault(m, {})
        try:
            a1, a2 = cc.update_amps(t1, t2, eris)
        finally:
            friccsd.TERM_PROBE = None
        outs[m] = (a1, a2, cc.energy(a1, a2, eris), cc._slots[0].indices())
    r1, r2, e_ref, i_ref = outs["fp64"]
    b1, b2, e_m, i_m = outs[mode]
    terms = {}
    for spec, ref in probes["fp64"].items():
        d = probes[mode][spec] - ref
        terms[spec] = {"max_rel": d.abs().max().item() / max(ref.abs().max().item(), 1e-300),
                       "fro_rel": (d.norm() / max(ref.norm().item(), 1e-300)).item()}
    res = {"same_sampled_T2": bool((i_ref == i_m).all()), "terms": terms,
           "t2_max_rel": (b2 - r2).abs().max().item() / r2.abs().max().item(),
           "t2_fro_rel": ((b2 - r2).norm() / r2.norm()).item(),
           "t1_max_rel": (b1 - r1).abs().max().item() / max(r1.abs().max().item(), 1e-300),
           "dE_corr": abs(e_m - e_ref), "E_corr_fp64": e_ref}
    del probes, outs
    torch.cuda.empty_cache()
    return res


def dense_work(eris, cfg, mode, t1, t2):
    """Algorithmic flops and bytes of one step's dense products, counted per
    product by fricc.dense (one step from (t1, t2))."""
    from fricc import dense
    from fricc.friccsd import FRICCSD
    from fricc.uniforms import FixedUniforms

    cc = FRICCSD(eris, {"m_keep": cfg["m_keep"], "compression": "fri", "sampling_method": "systematic",
                        "precision": mode}, uniforms=FixedUniforms([0.5, 0.5]))
    cc.level_shift = 2.0
    dense.COUNTER = {}
    try:
        cc.update_amps(t1, t2, eris)
        return dict(dense.COUNTER)
    finally:
        dense.COUNTER = None


def iteration_roofline(o, v, m, mode, dense_rec, ms_measured):
    """T_roof / T_measured of one iteration: T_roof sums max(F / P, B / BW)
    over the parts, with the algorithmic work of each (compression: two
    streaming passes of x and the COO output per sample; contractions:
    2 m (v^2 + 2 o^2 + 8 o v) flops in the mode's precision, the ladder's
    operand bytes v^4 s_W plus the FP64 output read and written; the dense
    products as counted; the finalize's o^2 v^2 arrays once each)."""
    N = o * o * v * v
    sp, lp = MODE_PREC[mode]
    peak = lambda p: 1e12 * PEAK_TFLOPS[p]  # noqa: E731
    bw = 1e9 * HBM_GBS
    npass = lambda p: 3 if p == "bf16" else 1  # noqa: E731
    w_bytes = {"fp64": 8, "fp32": 4, "bf16": 4, "bf16x1": 2}
    parts = {}
    parts["compression x2"] = 2 * (2 * 8 * N + 12 * m) / bw
    lf = 2.0 * m * v * v * npass(lp)
    parts["ladder 2323"] = max(lf / peak(lp.replace("x1", "")), (float(v) ** 4 * w_bytes[lp] + 2.0 * N * 8) / bw)
    of = 16.0 * m * o * v * npass(sp)
    parts["O3V3 (8 terms)"] = max(of / peak(sp), 8 * 3 * N * w_bytes[sp] / bw)
    parts["O4V2 (2 terms)"] = 4.0 * m * o * o * npass(sp) / peak(sp)
    tdense = 0.0
    for rec in dense_rec.values():
        tdense += max(rec["flops"] / peak("fp64" if rec["fp64"] else "fp32"), rec["bytes"] / bw)
    parts["dense remainder"] = tdense
    # ovovT, t2old and the direct terms read, t2new written (FP64), S read in
    # the mode's precision; the reduced modes also write the FP32 copy of t2new
    # their next step reads (Lovov is formed from ovovT, not read)
    reduced = mode != "fp64"
    parts["finalize"] = ((4 * 8 + (4 if reduced else 8) + (4 if reduced else 0)) * N) / bw
    t_roof = sum(parts.values())
    return {"T_roof_ms": 1e3 * t_roof, "T_measured_ms": ms_measured, "frac": 1e3 * t_roof / ms_measured,
            "parts_ms": {k: 1e3 * x for k, x in parts.items()},
            "dense_flops": sum(r["flops"] for r in dense_rec.values()),
            "dense_bytes": sum(r["bytes"] for r in dense_rec.values())}


# (W shape, output shape, number of output blocks the reference loops over)
# for each contraction term, src/sparse_tensor.cpp:11-346
def _term_geometry(o, v):
    ovov, vo = (o, v, o, v), v * o
    return {
        "0101": ((o, o, o, o), (o, o, v, v), v * v), "2323": ((v, v, v, v), (o, o, v, v), o * o),
        "1302": ((v, o, o, v), (o, o, v, v), vo), "1202": ((v, o, v, o), (o, o, v, v), vo),
        "1303": ((v, o, o, v), (o, o, v, v), vo), "1203": ((v, o, v, o), (o, o, v, v), vo),
        "1323": (ovov, (o, o, o, o), o * o), "0112": (ovov, (v, o, o, v), vo), "0313": (ovov, (v, o, o, v), vo),
        "0113": (ovov, (v, o, o, v), vo), "0312": (ovov, (v, o, v, o), vo),
    }


def cpu_baseline_sample(x, idx, val, o, v, m, blocks_per_term=None, slice_len=1 << 23, log=None):
    """CPU path of one FRI-CCSD iteration, timed on a bounded sample and
    scaled to seconds per iteration.  The timed code is the CPU oracle's
    restatement of the reference's native path (oracle/fricc_oracle.cpp):

    * compression: the oracle's FAITHFUL mode (sequential 1-norm, D loop and
      partial_sum, forward systematic scan) on the first ``slice_len`` entries
      of T2 with m scaled to the slice, times N/slice_len, twice per iteration
      (t2_sparse and t2_sparse_alt).  Its top-m selection is a partial sort;
      the reference sorts all N (src/fri_utils.cpp:11-36), so this part is
      optimistic for the CPU.
    * the 11 sparse contractions: the reference's loop nests over the same AoS
      entries (the iteration's real sampled T2), a contiguous range of output
      blocks from the middle of each term's block range, scaled by the block
      count (2323: by the entry count, its per-entry v^2 loop dominating).
      W is a constant array of the reference's shape and C layout (the values
      do not change the loop's work).
    * the dense remainder: 11 O(o^2 v^3) contractions (fricc/friccsd.py:144-207
      via PySCF), timed as 11 host DGEMMs of that size.

    Returns (seconds per iteration, threads, description)."""
    import numpy as np
    import torch

    from oracle import liboracle as lo

    say = log or (lambda *a: None)
    threads = lo.num_threads()
    N = x.size
    geo = _term_geometry(o, v)
    tdims = (o, o, v, v)
    bpt = blocks_per_term or 2 * threads
    # compression
    n_s = min(slice_len, N)
    m_s = max(1, int(round(m * n_s / N)))
    xs = np.ascontiguousarray(x[:n_s])
    lo.compress(xs[: min(n_s, 1 << 16)], max(1, m_s // 128), lo.FRI_SYSTEMATIC, 0.5, mode=lo.FAITHFUL)
    t0 = time.perf_counter()
    info = lo.compress(xs, m_s, lo.FRI_SYSTEMATIC, 0.5, mode=lo.FAITHFUL)[2]
    t_comp = (time.perf_counter() - t0) * N / n_s
    say(f"compression sample {n_s} entries, m={m_s}: {t_comp:.2f} s per call (status {info['status']})")
    # sparse contractions
    idx4 = np.ascontiguousarray(lo.unpack(idx, tdims))
    val = np.ascontiguousarray(val, dtype=np.float64)
    per_term = {}
    for term, (wshape, oshape, nb) in geo.items():
        W = torch.full(wshape, 1e-3, dtype=torch.float64).numpy()
        out = np.zeros(oshape)
        b = min(threads if (term == "2323" and blocks_per_term is None) else bpt, nb)  # 2323: one block per thread
        lo_b = (nb - b) // 2
        t0 = time.perf_counter()
        lo.contract_ref(term, W, idx, val, tdims, out, lo_b, lo_b + b, idx4=idx4)
        t = time.perf_counter() - t0
        if term == "2323":
            blk = idx4[:, 0] * o + idx4[:, 1]
            e_r = int(np.count_nonzero((blk >= lo_b) & (blk < lo_b + b)))
            scale = idx.size / max(e_r, 1)
        else:
            scale = nb / b
        per_term[term] = t * scale
        say(f"{term}: blocks [{lo_b}, {lo_b + b}) of {nb}: {t:.3f} s -> {per_term[term]:.1f} s")
        del W, out
    # dense remainder
    A = np.random.default_rng(0).random((v, v))
    B = np.random.default_rng(1).random((v, o * o * v))
    A @ B
    t0 = time.perf_counter()
    A @ B
    t_dense = 11 * (time.perf_counter() - t0)
    total = 2 * t_comp + sum(per_term.values()) + t_dense
    desc = (f"oracle FAITHFUL compression of {n_s} of {N} T2 entries (x2 per iteration, scaled by N); "
            f"reference loop nests over the iteration's {idx.size} sampled entries, {bpt} output blocks per "
            f"term ({threads} for 2323; scaled by block count, 2323 by entry count); dense remainder as 11 host DGEMMs "
            f"{v}x{v}x{o * o * v}; {threads} threads")
    return total, threads, desc


COPY_BYTES = 1 << 30  # calibration copy for the traffic counters


def replica_bench(args):
    """Batched replica trajectories of the o=20, v=180 system (naux 600,
    seed 1): R replicas with m log-uniform in [2e5, 2e6] and Philox seeds
    0..R-1, one graph launch per iteration for all of them, against the same
    replicas stepped one at a time eagerly."""
    import gc

    import torch

    from fricc.replicas import ReplicaBatch, log_uniform_m, run_alone
    from fricc.systems import CONFIGS, DeviceERIs, SyntheticSystem

    # blocks cached by earlier legs (the serving leg keeps retired solves'
    # buffers) go back to the device first: with the batch's 64 replicas and
    # arenas resident, allocations of the eager replicas would otherwise hit
    # the allocator's free-and-retry path
    gc.collect()
    torch.cuda.empty_cache()
    cfg = CONFIGS["c2"]
    o, v = cfg["nocc"], cfg["nvir"]
    sysm = SyntheticSystem.from_config("c2")
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc)
    R = args.replicas
    m_list = log_uniform_m(R, 2e5, 2e6, seed=args.seed, cap=o * o * v * v)
    seeds = list(range(R))
    settings = {"compression": "fri", "sampling_method": "systematic", "precision": args.replica_precision}
    batch = ReplicaBatch(eris, settings, m_list, seeds, level_shift=2.0)
    t0 = time.perf_counter()
    batch.step()  # records the graph
    t_record = time.perf_counter() - t0
    K = args.replica_iters
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        batch.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the same work stepped eagerly, one replica at a time (a sample of replicas)
    ns = min(4, R)
    run_alone(eris, settings, m_list[0], seeds[0], 1, level_shift=2.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(ns):
        run_alone(eris, settings, m_list[r], seeds[r], 2, level_shift=2.0)
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) / (2 * ns)
    s = batch.summary()
    return {"system": f"synthetic DF nocc={o} nvir={v} naux={cfg['naux']} seed={cfg['seed']}, level_shift=2.0",
            "replicas": R, "m_range": [min(m_list), max(m_list)], "precision": args.replica_precision,
            "timed_iterations": K + 1, "replica_iter_per_s": R * K / dt, "ms_per_batch_iteration": 1e3 * dt / K,
            "eager_replica_iter_per_s": 1.0 / eager, "record_s": t_record,
            "E_corr_mean": s["mean"], "E_corr_stderr": s["stderr"]}


def traffic_child(args):
    """Run under rocprofv3 --pmc by measure_traffic: one calibration copy of
    COPY_BYTES, then the headline mode's warm-up and two iterations of the
    timed window, whose longest ladder launch is read."""
    import ctypes

    import torch

    from fricc import _capi
    from fricc.systems import CONFIGS, DeviceERIs, SyntheticSystem

    torch.cuda.set_device(0)
    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    _capi.check(_capi.lib().fricc_stream_copy(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                              COPY_BYTES, _capi.stream_ptr()))
    torch.cuda.synchronize()
    del a, b
    cfg = CONFIGS[args.config]
    sysm = SyntheticSystem.from_config(args.config)
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc)
    run_mode(eris, cfg, args.headline, 2, args.warmup, args.seed)
    torch.cuda.synchronize()


def measure_traffic(args, head):
    """HBM bytes and MFMA use of one ladder launch from three rocprofv3 --pmc
    runs of traffic_child: FETCH_SIZE, then WRITE_SIZE (KiB per dispatch),
    each scaled by what the same counter reads for the COPY_BYTES copy (which
    moves exactly COPY_BYTES each way); then the derived MfmaUtil (MFMA-busy
    cycles over active cycles x SIMDs, %) and MfmaFlopsBF16 (BF16 MFMA flops
    issued).  Returns (dict or None, note)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    base = tempfile.mkdtemp(prefix="fricc_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    got = {}
    for counters in (("FETCH_SIZE",), ("WRITE_SIZE",), ("MfmaUtil", "MfmaFlopsBF16")):
        d = os.path.join(base, counters[0])
        cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               "-u", os.path.abspath(__file__), "--traffic-child", "--config", args.config, "--seed",
               str(args.seed), "--headline", head, "--warmup", str(args.warmup)]
        try:
            subprocess.run(cmd, timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=env,
                           cwd="/tmp", check=False)
        except subprocess.TimeoutExpired:
            return None, f"{counters[0]} pass timed out"
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None, f"{counters[0]} pass wrote no counters"
        rows = list(csv.DictReader(open(files[0])))
        for c in counters:
            crow = [r for r in rows if r["Counter_Name"] == c]
            copy = [float(r["Counter_Value"]) for r in crow if "k_stream_copy" in r["Kernel_Name"]]
            gem = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]), float(r["Counter_Value"]))
                   for r in crow if "k_ggemm_bf16" in r["Kernel_Name"]]
            if not gem:
                return None, f"{c} pass: ladder dispatch missing"
            # the ladder is the longest gathered launch of the window
            got[c] = (copy[0] if copy else None, max(gem)[1], max(gem)[0])
    shutil.rmtree(base, ignore_errors=True)
    if got["FETCH_SIZE"][0] is None or got["WRITE_SIZE"][0] is None:
        return None, "copy dispatch missing"
    f_fetch = COPY_BYTES / (got["FETCH_SIZE"][0] * 1024.0)
    f_write = COPY_BYTES / (got["WRITE_SIZE"][0] * 1024.0)
    fetch = got["FETCH_SIZE"][1] * 1024.0 * f_fetch
    write = got["WRITE_SIZE"][1] * 1024.0 * f_write
    return {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write, "fetch_scale": f_fetch,
            "write_scale": f_write, "mfma_util_pct": got["MfmaUtil"][1],
            "mfma_flops_issued": got["MfmaFlopsBF16"][1],
            "launch_ms_under_profiler": got["FETCH_SIZE"][2] / 1e6}, "ok"


def serve_bench(args):
    """Continuous-batching serving: Poisson arrivals (rate args.serve_rate per
    second) of args.serve solves, (nocc, nvir) uniform over the given sizes,
    m/|T2| uniform in [1%, 10%], each to |dE| < 1e-6 Ha or 50 iterations,
    admission by an HBM budget of 240 GB, mixed precision."""
    import numpy as np
    import torch

    from fricc import serving as sv

    sizes = [tuple(int(x) for x in s.split("x")) for s in args.serve_sizes.split(",")]
    reqs = sv.poisson_requests(args.serve, args.serve_rate, sizes, (0.01, 0.10), seed=args.seed)
    def progress(rec):
        t, ev, rid, o, v = rec[:5]
        print(f"# serve t={t:.2f}s {ev} {rid} ({o}x{v})", file=sys.stderr, flush=True)

    # up to 8 solves admitted, their iterations interleaved (shortest remaining
    # work first by default), one solve's work at a time on the device
    server = sv.Server(hbm_budget=240e9, precision="mixed", max_active=args.serve_active, progress=progress,
                       policy=args.serve_policy, overlap=args.serve_overlap)
    torch.cuda.empty_cache()
    at_start = {"allocated": torch.cuda.memory_allocated(), "device_free": torch.cuda.mem_get_info()[0]}
    print(f"# serving: memory at start {at_start}", file=sys.stderr, flush=True)
    res = server.run(reqs)
    st = sv.latency_stats(res, server.makespan)
    st.update({"sizes": [f"{o}x{v}" for o, v in sizes], "rate_per_s": args.serve_rate,
               "max_active": args.serve_active, "policy": args.serve_policy,
               "device_order": ("side by side (overlap)" if args.serve_overlap
                                else "one solve at a time (interleaved iterations)"),
               "max_concurrent": server.max_concurrent, "hbm_budget_gb": 240,
               "footprint_gb": {f"{o}x{v}": round(b / 1e9, 2) for (o, v), b in server.measured.items()},
               "memory_at_start": at_start, "admission_log": server.log[:64],
               "per_size": {f"{o}x{v}": {"n": sum(1 for r in res if (r.nocc, r.nvir) == (o, v)),
                                         "p50_s": float(np.median([r.latency for r in res if (r.nocc, r.nvir) == (o, v)]
                                                                  or [float("nan")]))}
                            for o, v in sizes}})
    torch.cuda.empty_cache()
    return st


def cpu_model():
    """The host CPU's model name and logical core count (lscpu / os)."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": name, "logical_cpus": os.cpu_count()}


def summarize_profile(path, only=None):
    """Markdown tables from a rocprofv3 --kernel-trace --stats output
    directory (its *kernel_stats.csv): per kernel the calls, total and average
    duration, the 30 longest by total time; then per-iteration tables."""
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(path, "**", "*kernel_stats.csv"), recursive=True))
    if not files:
        raise SystemExit(f"no *kernel_stats.csv under {path}")
    rows = list(csv.DictReader(open(files[0])))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:30]:
        name = r["Name"]
        name = name if len(name) <= 90 else name[:87] + "..."
        out.append(f"| `{name}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {100 * float(r['TotalDurationNs']) / tot:.1f} |")
    out.append(f"\nTotal kernel time {tot / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} dispatches "
               f"({os.path.basename(files[0])}).")
    out += step_breakdown(path, only=only)
    return "\n".join(out)


def step_breakdown(path, top=16, only=None):
    """Per-iteration kernel tables from the same directory's *kernel_trace.csv:
    an iteration is the dispatches after one k_t2_finalize (and its partial
    sums) up to and including the next one's partial sums -- the two
    compressions, the views, every contraction, the dense remainder and the
    fused finalize."""
    import collections
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True))
    if not files:
        return []
    rows = sorted(csv.DictReader(open(files[0])), key=lambda r: int(r["Start_Timestamp"]))
    fin = [i for i, r in enumerate(rows) if "k_t2_finalize" in r["Kernel_Name"]]
    out = []
    longest = {}
    for w in range(1, len(fin)):
        seg = rows[fin[w - 1] + 2:fin[w] + 2]
        lad = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg if "k_ggemm_bf16" in r["Kernel_Name"]]
        if lad:
            longest[w + 1] = max(lad) / 1e6
        wall = (int(seg[-1]["End_Timestamp"]) - int(seg[0]["Start_Timestamp"])) / 1e6
        tot = collections.defaultdict(lambda: [0, 0])
        for r in seg:
            k = r["Kernel_Name"].split("(")[0].replace("void ", "")
            k = k if len(k) <= 70 else k[:67] + "..."
            tot[k][0] += 1
            tot[k][1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        busy = sum(v[1] for v in tot.values()) / 1e6
        if only and (w + 1) not in only:
            continue
        out.append(f"\n### Iteration {w + 1} of the trace: {len(seg)} dispatches, kernel time {busy:.2f} ms, "
                   f"wall {wall:.2f} ms\n")
        out.append("| ms | calls | kernel |")
        out.append("|---:|---:|---|")
        for k, (cnt, d) in sorted(tot.items(), key=lambda x: -x[1][1])[:top]:
            out.append(f"| {d / 1e6:.3f} | {cnt} | `{k}` |")
    if longest:
        out.append("\n### Ladder launch (the longest `k_ggemm_bf16` of each iteration), ms\n")
        out.append(" ".join(f"{w}:{t:.3f}" for w, t in sorted(longest.items())))
        win = [t for w, t in longest.items() if 6 <= w <= 25]
        if win:
            out.append(f"\nMean over iterations 6..25 (the bench's timed window): {sum(win) / len(win):.3f} ms "
                       f"over {len(win)} launches.")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--iterations", type=int, default=40, help="iterations of the headline trajectory (c3: 40)")
    ap.add_argument("--config", default="c3")
    ap.add_argument("--modes", default="mixed,bf16-ladder,fp32,fp64")
    ap.add_argument("--headline", default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 traffic passes")
    ap.add_argument("--no-errors", action="store_true", help="skip the per-term precision errors")
    ap.add_argument("--replicas", type=int, default=0, help="also run R batched replicas of the o=20 v=180 system")
    ap.add_argument("--replica-iters", type=int, default=3)
    ap.add_argument("--replica-precision", default="mixed")
    ap.add_argument("--serve", type=int, default=0, help="also serve N solves arriving as a Poisson stream")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-active", type=int, default=8, help="solves admitted at once (interleaved)")
    ap.add_argument("--serve-policy", default="srpt", choices=("srpt", "rr"),
                    help="next iteration: shortest remaining work first, or round robin")
    ap.add_argument("--serve-sizes", default="10x90,20x180,32x256")
    ap.add_argument("--serve-overlap", action="store_true", help="solves' iterations side by side on the device")
    ap.add_argument("--lccsd-modes", default="mixed,fp64", help="FRI-LCCSD iterations timed beside ('' for none)")
    ap.add_argument("--lccsd-steps", type=int, default=3)
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--summarize-profile", default=None, help="markdown summary of a rocprofv3 output dir; no GPU")
    ap.add_argument("--profile-iterations", default=None, help="iterations to tabulate in the summary, e.g. 6,15,25")
    args = ap.parse_args()
    if args.summarize_profile:
        only = {int(x) for x in args.profile_iterations.split(",")} if args.profile_iterations else None
        print(summarize_profile(args.summarize_profile, only))
        return
    if args.traffic_child:
        traffic_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path runs on one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from fricc.systems import CONFIGS, DeviceERIs, SyntheticSystem

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    cfg = CONFIGS[args.config]
    o, v = cfg["nocc"], cfg["nvir"]
    m = cfg["m_keep"]
    N = o * o * v * v
    t0 = time.perf_counter()
    sysm = SyntheticSystem.from_config(args.config)
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    modes = args.modes.split(",")
    head = args.headline if args.headline in modes else modes[0]
    res = {}
    for mode in [head] + [x for x in modes if x != head]:
        res[mode] = run_mode(eris, cfg, mode, args.steps, args.warmup, args.seed,
                             total=args.iterations if mode == head else None)
        print(f"# {mode}: " + json.dumps({k: v_ for k, v_ in res[mode].items()
                                          if k not in ("energies", "state", "slots", "latency_ms")}),
              file=sys.stderr, flush=True)
    lccsd = {}
    if args.lccsd_modes:
        from fricc.frilccsd import FRILCCSD

        for mode in args.lccsd_modes.split(","):
            rl = run_mode(eris, cfg, mode, args.lccsd_steps, 1, args.seed, solver=FRILCCSD)
            lccsd[mode] = {"value": rl["it_per_s"], "ms_per_step": rl["ms_per_step"], "p50_ms": rl["p50_ms"],
                           "E_corr_last": rl["energies"][-1]}
            del rl
            print(f"# FRI-LCCSD {mode}: {lccsd[mode]}", file=sys.stderr, flush=True)
    r = res[head]
    t1s, t2s = r["state"]
    # per-term errors of every reduced mode against FP64 on the same sampled T2
    errs = {}
    if not args.no_errors:
        for mode in res:
            if mode != "fp64":
                errs[mode] = per_term_errors(eris, cfg, mode, t1s, t2s)
                print(f"# error {mode}: {json.dumps(errs[mode])}", file=sys.stderr, flush=True)
    # whole-iteration roofline of every mode (dense work counted in one step)
    whole = {}
    for mode in res:
        rec = dense_work(eris, cfg, mode, t1s, t2s)
        whole[mode] = iteration_roofline(o, v, m, mode, rec, res[mode]["ms_per_step"])
    sp, lp = MODE_PREC[head]
    lad_ms = r["sections_ms"].get("ladder")
    # ladder 2323, one launch: 2 m v^2 flops; the v^4 operand in its stored
    # form plus the FP64 output read and written
    w_bytes = {"fp64": 8, "fp32": 4, "bf16": 4, "bf16x1": 2}[lp]
    lad_flops = 2.0 * m * v * v
    lad_bytes = float(v) ** 4 * w_bytes + 2.0 * N * 8
    # achievable HBM bandwidth: the streaming copy kernel, HIP events on its stream
    import ctypes

    from fricc import _capi

    ca = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    cb = torch.empty_like(ca)
    cp = lambda: _capi.check(_capi.lib().fricc_stream_copy(ctypes.c_void_p(ca.data_ptr()),  # noqa: E731
                                                           ctypes.c_void_p(cb.data_ptr()), COPY_BYTES,
                                                           _capi.stream_ptr()))
    cp()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        cp()
    ev1.record()
    torch.cuda.synchronize()
    copy_gbs = 2.0 * COPY_BYTES * 10 / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del ca, cb
    if lp.startswith("bf16"):
        roof = {"kernel": "gathered ladder GEMM 2323 (k_ggemm_bf16_pl)", "bound": "hbm",
                "achieved": lad_bytes / (lad_ms * 1e-3) / 1e9 if lad_ms else None, "peak": HBM_GBS, "unit": "GB/s"}
    else:
        roof = {"kernel": "gathered ladder GEMM 2323 (k_ggemm)", "bound": "mfma",
                "achieved": lad_flops / (lad_ms * 1e-3) / 1e12 if lad_ms else None, "peak": PEAK_TFLOPS[lp],
                "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"] if roof["achieved"] else None
    roof["traffic"] = None
    roof.update({"launch_ms": lad_ms, "algorithmic_bytes_per_launch": lad_bytes,
                 "algorithmic_flops_per_launch": lad_flops,
                 "algorithmic_tflops": lad_flops / (lad_ms * 1e-3) / 1e12 if lad_ms else None,
                 "achievable_hbm_gbs_copy": copy_gbs, "timed_window": f"iterations {args.warmup + 1}.."
                                                                     f"{args.warmup + args.steps}"})
    lat = r["latency_ms"]
    out = {
        "metric": "FRI-CCSD iter/s at o=32,v=256,m=4e6 (fp32/bf16 vs fp64); p50/p99 solve latency",
        "value": r["it_per_s"],
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16-ladder": "bf16/fp32", "mixed": "bf16x3/fp32", "fp32": "fp32", "fp64": "fp64"}[head],
        "data": "synthetic",
        "config": {"workload": f"FRI-CCSD iteration, synthetic DF system nocc={o} nvir={v} naux={cfg['naux']} "
                               f"seed={cfg['seed']}, m_keep={m}, systematic FRI, level_shift=2.0",
                   "nocc": o, "nvir": v, "m_keep": m, "precision_mode": head},
        "iteration_latency_ms": {"p50": r["p50_ms"], "p99": r["p99_ms"]},
        "trajectory": {"iterations": len(lat), "latency_ms": [round(x, 3) for x in lat],
                       "mean_ms": sum(lat) / len(lat), "it_per_s": 1e3 * len(lat) / sum(lat),
                       "E_corr": r["energies"]},
        "sections_ms": r["sections_ms"],
        "roofline": roof,
        "iteration_roofline": whole,
        "setup_s": t_setup,
        "n_D": r["n_D"],
        "n_sampled": r["n_sampled"],
    }
    out["modes"] = {m_: {"value": res[m_]["it_per_s"], "ms_per_step": res[m_]["ms_per_step"],
                         "p50_ms": res[m_]["p50_ms"], "p99_ms": res[m_]["p99_ms"],
                         "sections_ms": res[m_]["sections_ms"]} for m_ in res}
    if "fp64" in res and head != "fp64":
        out["speedup_vs_fp64"] = r["it_per_s"] / res["fp64"]["it_per_s"]
    out["precision_error_vs_fp64"] = errs
    if lccsd:
        out["frilccsd"] = {"unit": "iter/s", "modes": lccsd,
                           "workload": "FRI-LCCSD iteration on the same system, m and uniforms"}
    out["cpu_baseline"] = None
    # host copies for the CPU leg, then free the GPU for the profiling passes
    t2_host = t2s.cpu().numpy().ravel()
    ts = r["slots"][0]
    ts_idx, ts_val = ts.indices(), ts.values()
    for k in list(res):
        res[k].pop("state", None)
        res[k].pop("slots", None)
    del eris, ts, sysm
    t1s = t2s = None
    import gc

    gc.collect()
    torch.cuda.empty_cache()
    print(f"# memory after the timed modes: allocated {torch.cuda.memory_allocated() / 1e9:.2f} GB, "
          f"device free {torch.cuda.mem_get_info()[0] / 1e9:.2f} GB", file=sys.stderr, flush=True)
    if not args.no_traffic:
        tr, note = measure_traffic(args, head)
        print(f"# traffic: {tr} ({note})", file=sys.stderr, flush=True)
        if tr is not None:
            roof["traffic"] = tr["bytes"]
            roof["traffic_detail"] = tr
        else:
            roof["traffic_note"] = note
    if args.serve > 0:
        out["serving"] = serve_bench(args)
        sv_ = out["serving"]
        out["solve_latency_s"] = {"p50": sv_["p50_s"], "p99": sv_["p99_s"], "solves_per_s": sv_["solves_per_s"],
                                  "workload": f"{sv_['solves']} synthetic DF solves, Poisson arrivals at "
                                              f"{sv_['rate_per_s']}/s, sizes {sv_['sizes']}, m/|T2| in [1%, 10%], "
                                              f"each to |dE| < 1e-6 Ha or 50 iterations, mixed precision"}
        print(f"# serving: {out['serving']}", file=sys.stderr, flush=True)
    if args.replicas > 0:
        out["replica_batch"] = replica_bench(args)
        print(f"# replicas: {out['replica_batch']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        t_cpu, thr, desc = cpu_baseline_sample(
            t2_host, ts_idx, ts_val, o, v, m, log=lambda s: print(f"# cpu: {s}", file=sys.stderr, flush=True))
        cpu = cpu_model()
        out["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "iter/s", "cores": thr, "kind": "port",
                               "s_per_iter": t_cpu, "cpu": cpu["model"], "logical_cpus": cpu["logical_cpus"],
                               "sample": desc}
        out["speedup_vs_cpu"] = r["it_per_s"] * t_cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
