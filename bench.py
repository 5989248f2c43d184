"""FRI-CCSD iterations/s on one MI355X (GPU) at o=32, v=256, m=4e6.

One step = one FRI-CCSD iteration of the reference's kernel loop
(fricc/friccsd.py:416-430): update_amps (two FRI compressions of T2, the 11
sparse contractions plus the ladder's t1 dressing, the dense remainder,
denominators), the amplitude-change norm and the energy.  Workload: the
synthetic density-fitted system (nocc=32, nvir=256, naux=864, seed 1;
fricc.systems), MP2 guess, systematic FRI with m = 4e6 uniforms from
srand(1)/rand(), level shift 2.0 (plain Jacobi diverges on this system), all
state resident in HBM before the timed region.

The headline value is the precision mode ("mixed": FP32 sparse contractions
and dense GEMMs; FP64 compression, energy, denominators); the FP64 plain form
is timed beside it.  Prints one JSON line.
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


def _pct(a, q):
    import numpy as np

    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def run_mode(eris, cfg, mode, steps, warmup, seed):
    import torch

    from fricc.friccsd import FRICCSD, SectionTimer
    from fricc.uniforms import GlibcUniforms

    settings = {"m_keep": cfg["m_keep"], "compression": "fri", "sampling_method": "systematic",
                "precision": mode}
    cc = FRICCSD(eris, settings, uniforms=GlibcUniforms(seed))
    cc.level_shift = 2.0
    t1, t2 = cc.get_init_guess(eris)
    e = cc.energy(t1, t2, eris)

    def step(t1, t2, e):
        # as the kernel loop: the energy and the amplitude-change norm come
        # from the fused finalize pass of update_amps (read back once)
        t1n, t2n = cc.update_amps(t1, t2, eris)
        _, _, en, normt = cc._last_step
        return t1n, t2n, en, normt

    for _ in range(warmup):
        t1, t2, e, _ = step(t1, t2, e)
    torch.cuda.synchronize()
    cc.timer = SectionTimer()
    lat = []
    energies = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        t1, t2, e, normt = step(t1, t2, e)  # energy read back each cycle, as the loop does
        lat.append(time.perf_counter() - ts)
        energies.append(e)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sec = cc.timer.ms()
    info = cc.sparse_info[0]
    cc.final_amplitudes = (t1, t2)
    return {"state": (t1, t2), "it_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps, "p50_ms": 1e3 * _pct(lat, 50),
            "p99_ms": 1e3 * _pct(lat, 99), "sections_ms": {k: sum(v) / len(v) for k, v in sec.items() if v},
            "energies": energies, "n_D": info["n_D"], "n_sampled": info["n_sampled"]}


def precision_error(eris, cfg, mode, t1, t2, seed):
    """One step from the same (t1, t2) with the same uniforms in FP64 and in
    ``mode``: both compress the same T2 (compression is FP64 in every mode),
    so the difference is the arithmetic alone."""
    import torch

    from fricc.friccsd import FRICCSD
    from fricc.uniforms import FixedUniforms

    us = [0.3141592653589793, 0.2718281828459045]
    out = {}
    settings = {"m_keep": cfg["m_keep"], "compression": "fri", "sampling_method": "systematic"}
    c64 = FRICCSD(eris, dict(settings, precision="fp64"), uniforms=FixedUniforms(us))
    cm = FRICCSD(eris, dict(settings, precision=mode), uniforms=FixedUniforms(us))
    c64.level_shift = cm.level_shift = 2.0
    a1, a2 = c64.update_amps(t1, t2, eris)
    b1, b2 = cm.update_amps(t1, t2, eris)
    out["t2_max_rel"] = (b2 - a2).abs().max().item() / a2.abs().max().item()
    out["t1_max_rel"] = (b1 - a1).abs().max().item() / max(a1.abs().max().item(), 1e-300)
    out["dE_corr"] = abs(c64.energy(a1, a2, eris) - cm.energy(b1, b2, eris))
    out["E_corr_fp64"] = c64.energy(a1, a2, eris)
    del a1, a2, b1, b2
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--modes", default="mixed,fp32,fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
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
    t0 = time.perf_counter()
    sysm = SyntheticSystem.from_config(args.config)
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    res = {}
    for mode in args.modes.split(","):
        res[mode] = run_mode(eris, cfg, mode, args.steps, args.warmup, args.seed)
        print(f"# {mode}: {json.dumps({k: v for k, v in res[mode].items() if k not in ('energies', 'state')})}",
              file=sys.stderr, flush=True)
    errs = {}
    if "fp64" in res:
        t1s, t2s = res[list(res)[0]]["state"]
        for mode in res:
            if mode != "fp64":
                errs[mode] = precision_error(eris, cfg, mode, t1s, t2s, args.seed)
                print(f"# error {mode}: {errs[mode]}", file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in res else list(res)[0]
    r = res[head]
    ladder_flops = 2.0 * m * v * v
    lad_ms = r["sections_ms"].get("ladder")
    prec = {"mixed": "bf16", "fp32": "fp32", "fp64": "fp64"}[head]
    achieved = ladder_flops / (lad_ms * 1e-3) / 1e12 if lad_ms else None
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
        "dtype": {"mixed": "bf16/fp32", "fp32": "fp32", "fp64": "fp64"}[head],
        "data": "synthetic",
        "config": {"workload": f"FRI-CCSD iteration, synthetic DF system nocc={o} nvir={v} naux={cfg['naux']} "
                               f"seed={cfg['seed']}, m_keep={m}, systematic FRI, level_shift=2.0",
                   "nocc": o, "nvir": v, "m_keep": m, "precision_mode": head},
        "iteration_latency_ms": {"p50": r["p50_ms"], "p99": r["p99_ms"]},
        "sections_ms": r["sections_ms"],
        "roofline": {"kernel": "gathered ladder GEMM (2323)", "bound": "mfma",
                     "achieved": achieved, "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                     "frac": (achieved / PEAK_TFLOPS[prec]) if achieved else None, "traffic": None,
                     "algorithmic_flops_per_launch": ladder_flops},
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
    out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
