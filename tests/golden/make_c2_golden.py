"""Writes tests/golden/c2.npz: digests of the oracle's FRI-CCSD trajectory on
the c2 system (synthetic DF, nocc=20, nvir=180, naux=600, seed 1, m = 1e6,
systematic FRI, uniforms u = rand()/RAND_MAX after srand(1), level shift 2.0,
all contractions compressed), for the GPU parity tests at that size.

Run in a container with the CPU oracle built (make -C oracle); takes several
minutes and ~30 GB of memory (the v^4 integrals and W_vvvv).  The oracle is
the NumPy/C++ restatement of the reference (oracle/rccsd.py, oracle/
fricc_oracle.cpp); nothing of the reference itself is run.

Digests per iteration k (inputs of iteration 1 = the MP2 guess):
  E[k], normt[k]                        energy and amplitude-change norm
  idx_sha[k, s], val_sha[k, s]          SHA-256 of the kept flat indices
                                        (uint32 LE) and values (float64 LE) of
                                        sample s (0: t2_sparse, 1: t2_sparse_alt)
  val_at[k, s]                          kept values at VAL_POSITIONS (after
                                        iteration 1 the compressed T2 carries
                                        the last-bit differences of the FP64
                                        update, so values compare to 1e-10)
  nD[k, s], nsamp[k, s]                 |D| and n' of each sample
  t1[k]                                 t1 after the iteration (o x v)
  t2_at[k]                              t2 after the iteration at POSITIONS
  t2_sum[k], t2_sumsq[k], t2_wsum[k]    checksums of t2 (w from WEIGHT_SEED)
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import rccsd  # noqa: E402

NOCC, NVIR, NAUX, SEED, M = 20, 180, 600, 1, 1_000_000
LEVEL_SHIFT = 2.0
N_ITER = 3
POS_SEED, N_POS, WEIGHT_SEED = 2024, 20000, 7
VAL_SEED, N_VAL = 99, 4096


def positions(n):
    return np.sort(np.random.default_rng(POS_SEED).choice(n, N_POS, replace=False))


def val_positions(m):
    return np.sort(np.random.default_rng(VAL_SEED).choice(m, N_VAL, replace=False))


def weights(n):
    return np.random.default_rng(WEIGHT_SEED).standard_normal(n)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2.npz")):
    t0 = time.time()
    eris, _ = rccsd.synthetic_eris_blocks(NOCC, NVIR, NAUX, SEED)
    print(f"integrals {time.time() - t0:.1f} s", flush=True)
    emp2, t1, t2 = rccsd.init_amps(eris)
    N = t2.size
    pos, w = positions(N), weights(N)
    us = [u for _, u in zip(range(2 * N_ITER), rccsd.GlibcUniforms(SEED))]
    s = rccsd.default_settings(m_keep=M, level_shift=LEVEL_SHIFT)
    rec = {k: [] for k in ("E", "normt", "idx_sha", "val_sha", "val_at", "nD", "nsamp", "t1", "t2_at", "t2_sum",
                           "t2_sumsq", "t2_wsum")}
    vpos = val_positions(M)
    uit = iter(us)
    ecc = rccsd.energy(t1, t2, eris)
    for k in range(N_ITER):
        t1n, t2n, (ts, tsa) = rccsd.update_amps(t1, t2, eris, s, uit, return_sparse=True)
        normt = np.linalg.norm(rccsd.amplitudes_to_vector(t1n, t2n) - rccsd.amplitudes_to_vector(t1, t2))
        t1, t2 = t1n, t2n
        ecc = rccsd.energy(t1, t2, eris)
        rec["E"].append(ecc)
        rec["normt"].append(normt)
        rec["idx_sha"].append([sha(x.idx.astype("<u4")) for x in (ts, tsa)])
        rec["val_sha"].append([sha(x.val.astype("<f8")) for x in (ts, tsa)])
        rec["val_at"].append([x.val[vpos].copy() for x in (ts, tsa)])
        rec["nD"].append([x.info["n_D"] for x in (ts, tsa)])
        rec["nsamp"].append([x.info["n_sampled"] for x in (ts, tsa)])
        rec["t1"].append(t1.copy())
        flat = t2.ravel()
        rec["t2_at"].append(flat[pos].copy())
        rec["t2_sum"].append(flat.sum())
        rec["t2_sumsq"].append((flat * flat).sum())
        rec["t2_wsum"].append((flat * w).sum())
        print(f"iteration {k + 1}: E = {ecc:.15g} normt = {normt:.6g} |D| = {rec['nD'][-1]} "
              f"({time.time() - t0:.1f} s)", flush=True)
    np.savez_compressed(
        out, nocc=NOCC, nvir=NVIR, naux=NAUX, seed=SEED, m_keep=M, level_shift=LEVEL_SHIFT, uniforms=np.array(us),
        emp2=emp2, mp2_t2_sha=sha(rccsd.init_amps(eris)[2].astype("<f8")), pos_seed=POS_SEED, n_pos=N_POS,
        weight_seed=WEIGHT_SEED, val_seed=VAL_SEED, n_val=N_VAL, **{k: np.array(v) for k, v in rec.items()})
    print(f"wrote {out} ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main(*sys.argv[1:])
