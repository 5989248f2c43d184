"""Writes tests/golden/c3_compress.json: the oracle's compression of the c3
MP2 amplitudes (synthetic DF, nocc=32, nvir=256, naux=864, seed 1; N =
67,108,864, m = 4e6), two systematic samples with u = rand()/RAND_MAX after
srand(1) (t2_sparse, then t2_sparse_alt), as SHA-256 digests of the kept flat
indices (uint32 LE) and values (float64 LE) plus the compression statistics,
and a digest of the MP2 amplitudes themselves.  Also the same for "largest".

Run in a container with the CPU oracle built (make -C oracle); about a
minute.  Only the ovov block of the integrals is built (the MP2 guess needs
nothing else); each integral is the same fma chain as on the GPU."""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import liboracle, rccsd  # noqa: E402

NOCC, NVIR, NAUX, SEED, M = 32, 256, 864, 1, 4_000_000


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_compress.json")):
    t0 = time.time()
    L, e = rccsd.synthetic_L(NOCC, NVIR, NAUX, SEED)
    o, nmo = NOCC, NOCC + NVIR
    ovov = liboracle.eri_block(L, (0, o), (o, nmo), (0, o), (o, nmo))
    del L
    eia = e[:o, None] - e[None, o:]
    t2 = np.ascontiguousarray(ovov.transpose(0, 2, 1, 3) / (eia[:, None, :, None] + eia[None, :, None, :]))
    del ovov
    print(f"MP2 amplitudes {time.time() - t0:.1f} s", flush=True)
    g = rccsd.GlibcUniforms(SEED)
    us = [next(g), next(g)]
    rec = {"nocc": NOCC, "nvir": NVIR, "naux": NAUX, "seed": SEED, "m_keep": M, "uniforms": us,
           "mp2_t2_sha": sha(t2.astype("<f8")), "samples": []}
    x = t2.ravel()
    for u in us:
        idx, val, info = liboracle.compress(x, M, liboracle.FRI_SYSTEMATIC, u)
        assert info["status"] == 0, info
        rec["samples"].append({"u": u, "idx_sha": sha(idx.astype("<u4")), "val_sha": sha(val.astype("<f8")),
                               "n_D": int(info["n_D"]), "n_sampled": int(info["n_sampled"]),
                               "n_notfound": int(info["n_notfound"]), "norm0": float(info["norm0"]),
                               "remaining_norm": float(info["remaining_norm"])})
        print(f"sample u={u}: |D|={info['n_D']} ({time.time() - t0:.1f} s)", flush=True)
    idx, val, info = liboracle.compress(x, M, liboracle.LARGEST, 0.0)
    rec["largest"] = {"idx_sha": sha(idx.astype("<u4")), "val_sha": sha(val.astype("<f8"))}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(f"wrote {out} ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main(*sys.argv[1:])
