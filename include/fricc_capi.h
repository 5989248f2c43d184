/*
 * fricc_capi.h -- C ABI of libfricc_hip.so, the MI355X (gfx950) FRI-CCSD hot path.
 *
 * This is the drop-in boundary that replaces the reference's pybind11 module
 * `py_rccsd` (src/py_rccsd.cpp:16-55).  Plain C types only: device pointers
 * are raw `double*` / `uint32_t*`, shapes and strides are `int64_t[4]` in
 * elements, streams are `void*` (a hipStream_t; NULL = default stream).
 * Every function returns an int status (FRICC_OK == 0); on failure
 * fricc_last_error() describes it.  No function ever calls exit(): the cases
 * where the reference terminates the process (bad compression/term string,
 * p > 1, out-of-range index) return an error status instead.
 *
 * Threading: calls that share a workspace arena on one device are serialised
 * by the caller; the library keeps per-device scratch arenas, selected per
 * host thread (fricc_workspace_select).
 */
#ifndef FRICC_CAPI_H
#define FRICC_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRICC_ABI_VERSION 1

enum fricc_status {
  FRICC_OK = 0,
  FRICC_EINVAL = 1,     /* bad argument (reference: exit or pybind TypeError)   */
  FRICC_EDEVICE = 2,    /* HIP runtime error                                     */
  FRICC_ENOMEM = 3,     /* device allocation failed                              */
  FRICC_EPROB = 4,      /* a sampling probability exceeds 1 (src/fri_utils.cpp:383-395) */
  FRICC_ENONFINITE = 5, /* non-finite value in the vector to compress           */
  FRICC_ESAMPLE = 6     /* pivotal sampling ended with a sample count other than
                           n' (reference exits, src/fri_utils.cpp:100-123)       */
};

/* FRICC_FRI_PIVOTAL: "fri" with sampling_method "pivotal" (the binding's
 * default, src/py_rccsd.cpp:30,39).  Its chain draws one uniform per step
 * (rand()/RAND_MAX in the reference, src/fri_utils.cpp:73); here step j draws
 * the j-th output of SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) seeded
 * with the IEEE-754 bits of the call's uniform u, so one u per call still
 * fixes the sample. */
enum fricc_compression { FRICC_LARGEST = 0, FRICC_FRI_SYSTEMATIC = 1, FRICC_FRI_PIVOTAL = 2 };
/* FRICC_BF16: BF16 hi and lo planes of each operand, three MFMA passes
 * (hi*hi + hi*lo + lo*hi, about 16 significant bits per product);
 * FRICC_BF16X1: the hi plane only, one pass (BF16 products, FP32 accumulate). */
enum fricc_precision { FRICC_FP64 = 0, FRICC_FP32 = 1, FRICC_BF16 = 2, FRICC_BF16X1 = 3 };

typedef struct fricc_sparse fricc_sparse; /* device-resident SparseTensor4d */

typedef struct fricc_compress_info {
  int64_t n_D;            /* exactly kept entries |D|                         */
  int64_t n_sampled;      /* n' = m - |D|                                      */
  int64_t n_notfound;     /* samples past the last interval (index 0)          */
  int64_t n_candidates;   /* elements that entered the top-m sort              */
  double norm0;           /* 1-norm of x                                       */
  double remaining_norm;  /* 1-norm left after D (denominator of p)            */
  double p_total;         /* last interval end                                 */
  int32_t nonfinite;
  int32_t pad;
} fricc_compress_info;

int fricc_abi_version(void);
const char* fricc_last_error(void);

/* FRI compression of a device vector x[n] to m entries.
 * Replaces fri_compression(x, m, "systematic" | "pivotal")
 * (src/fri_utils.cpp:439-522, include/fri_utils.hpp:30-32) and, for
 * compression == FRICC_LARGEST, the argsort-top-m path of the SparseTensor4d
 * ctor (include/sparse_tensor.hpp:46-56).
 * `u` is the one uniform the reference draws per systematic call as
 * rand()/RAND_MAX (src/fri_utils.cpp:47); for FRICC_FRI_PIVOTAL it keys the
 * chain's step draws; it is ignored for FRICC_LARGEST.
 * out_idx/out_val: device arrays of m; flat indices, D first (|x| descending,
 * ties by index), then the samples (systematic: in ascending order; pivotal:
 * in the order the chain emits them, all distinct).  n < 2^31. */
int fricc_compress(const double* x, int64_t n, int64_t m, int compression, double u, uint32_t* out_idx,
                   double* out_val, fricc_compress_info* info, void* stream);

/* fricc_compress with no allocation of any kind, for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments, results and errors as fricc_compress; the scratch is a fixed
 * device array of the library (about 110 MB per GPU, held from the first
 * launch of any library kernel on that GPU until the process ends), so n is
 * limited to fricc_compress_noalloc_max_n() = 4194304 elements (FRICC_EINVAL
 * beyond).  All calls on one GPU share that array: calls that overlap on one
 * GPU (different streams or threads) race and must be serialised by the
 * caller.  It waits for the stream at the end, like fricc_compress. */
int fricc_compress_noalloc(const double* x, int64_t n, int64_t m, int compression, double u, uint32_t* out_idx,
                           double* out_val, fricc_compress_info* info, void* stream);
int64_t fricc_compress_noalloc_max_n(void);

/* sample_systematic(n_sample, probs) (src/fri_utils.cpp:129-159) on device
 * probabilities p[n] (each in [0,1]): S_k = j with P_j <= fl(k+u) < P_{j+1},
 * P the canonical (exact, once-rounded) interval ends; k past the last
 * interval keep index 0 (src/fri_utils.cpp:132).  out_idx: device array of
 * n_sample.  *n_notfound (host) receives the number of such k. */
int fricc_sample_systematic(const double* p, int64_t n, int64_t n_sample, double u, uint32_t* out_idx,
                            int64_t* n_notfound, void* stream);

/* sample_pivotal(n_sample, probs) (src/fri_utils.cpp:63-127) on device
 * probabilities p[n] (each in [0,1], else FRICC_EPROB), with the chain's mass
 * taken from the exact prefix of p (canonical: c_j = min(j, floor(Q_{j+1}))
 * units completed through step j, Q the exact sum of floor(p * 2^96) * 2^-96,
 * mass a_j = round(Q_{j+1} - c_j)).  Step j (1 <= j < n) draws uniforms[j]
 * (device array of n doubles; uniforms[0] is not read) or, with uniforms
 * NULL, the SplitMix64 stream keyed by the bits of u.  out_idx: device array
 * of n_sample, in the order the chain emits.  FRICC_ESAMPLE where the
 * reference exits (the chain ends with a count other than n_sample). */
int fricc_sample_pivotal(const double* p, int64_t n, int64_t n_sample, const double* uniforms, double u,
                         uint32_t* out_idx, void* stream);

/* parallel_sample(n_sample, probs) (src/fri_utils.cpp:176-362) with n_chunks
 * chunks in place of the reference's OpenMP threads (its result depends on
 * that count): chunk j = [j*step, (j+1)*step), the last one taking the
 * remainder, step = n / n_chunks.  Expected counts a_j = n_sample |p_j| / |p|
 * (canonical exact sums, rounded once), a pivotal chain over their fractional
 * parts picks the chunks that get one more sample, each chunk's
 * probabilities move in index order toward that count (up: min(1, p/t), down:
 * max(0, (p - t)/(1 - t)), exact running sum, the crossing element takes the
 * remainder), then each chunk runs the canonical pivotal chain; chunk j's
 * samples follow those of chunks 0..j-1.  Draws by position: the chunk
 * chain's step j takes uniforms[j*step], chunk j's step i uniforms[j*step+i]
 * (or the SplitMix64 stream keyed by u at those positions).  n_chunks in
 * [1, min(n, 65536)].  FRICC_EPROB / FRICC_ESAMPLE as fricc_sample_pivotal. */
int fricc_parallel_sample(const double* p, int64_t n, int64_t n_sample, int64_t n_chunks, const double* uniforms,
                          double u, uint32_t* out_idx, void* stream);

/* SparseTensor4d(dims, nnz) -- empty tensor (include/sparse_tensor.hpp:27-31). */
int fricc_sparse_create(const int64_t dims[4], fricc_sparse** out);
int fricc_sparse_destroy(fricc_sparse* sp);

/* SparseTensor4d(tensor_arr, dims, m, compression, "systematic")
 * (include/sparse_tensor.hpp:33-77): compress the device tensor x (C order,
 * prod(dims) doubles) into sp, reusing its storage. */
int fricc_sparse_compress(fricc_sparse* sp, const double* x, int64_t m, int compression, double u, void* stream);

/* The same compression with no host round trip, for a device-resident loop
 * (and graph capture): the uniform is read from device memory (u_dev, one
 * double, in [0, 1]); statistics and errors (non-finite input, p > 1, u out of
 * range) stay on the device until fricc_sparse_status.  No counterpart in the
 * reference. */
int fricc_sparse_compress_async(fricc_sparse* sp, const double* x, int64_t m, int compression, const double* u_dev,
                                void* stream);

/* Two samples of one vector: fricc_sparse_compress_async of x into sp with
 * the uniform at u_dev, then a second sample into sp2 with the uniform at
 * u_dev2.  The selection of D, the remaining norm and the prefix sums of p do
 * not depend on the uniform and are computed once; each tensor's entries and
 * statistics are bit for bit those of its own fricc_sparse_compress_async.
 * Replaces the two constructions of the same T2 in update_amps,
 * t2_sparse and t2_sparse_alt (fricc/friccsd.py:99-138). */
int fricc_sparse_compress_pair_async(fricc_sparse* sp, fricc_sparse* sp2, const double* x, int64_t m, int compression,
                                     const double* u_dev, const double* u_dev2, void* stream);

/* The same for count samples: sps[k] with the uniform at u_dev[k] (count
 * consecutive device doubles).  Replaces the per-term constructions of the
 * same T2 of FRILCCSD's "Independent Compressions" (fricc/frilccsd.py:233-457). */
int fricc_sparse_compress_multi_async(fricc_sparse* const* sps, int count, const double* x, int64_t m,
                                      int compression, const double* u_dev, void* stream);

/* Waits for work on `stream`, reads the scalars the last asynchronous
 * compression of sp left on the device (rewritten by every replay of a
 * recorded graph), makes its statistics visible to fricc_sparse_info and
 * returns its status (FRICC_OK, FRICC_ENONFINITE, FRICC_EPROB, or FRICC_EINVAL
 * for u outside [0, 1]).  FRICC_OK when sp was last filled synchronously. */
int fricc_sparse_status(fricc_sparse* sp, void* stream);

/* set_element for all entries at once (include/sparse_tensor.hpp:127-136):
 * host arrays of nnz flat indices and values; indices must be < prod(dims). */
int fricc_sparse_set(fricc_sparse* sp, const int64_t* idx, const double* val, int64_t nnz, void* stream);

/* dimensions / size / compression statistics */
int fricc_sparse_info(const fricc_sparse* sp, int64_t dims[4], int64_t* nnz, fricc_compress_info* info);

/* get_element for all entries (include/sparse_tensor.hpp:138-143): copies the
 * flat indices and values to host arrays of nnz. */
int fricc_sparse_get(const fricc_sparse* sp, int64_t* idx, double* val, void* stream);

/* device pointers of the entries (valid until the next compress/set/destroy) */
int fricc_sparse_device(const fricc_sparse* sp, const uint32_t** idx, const double** val);

/* contract_DTSpT(W, T, ouput, term) (src/sparse_tensor.cpp:349-382,
 * src/py_rccsd.cpp:50-51): out += term(W, T), accumulated in place.
 * term in {"0101","2323","1302","1202","1303","1203","1323","0112","0113",
 * "0313","0312"}; W and out are device FP64 4-D arrays with arbitrary strides
 * (in elements).  precision: FRICC_FP64, FRICC_FP32 (FP32 operands, FP32
 * accumulate), FRICC_BF16 (BF16 hi/lo operands, three MFMA passes, FP32
 * accumulate) or FRICC_BF16X1 (BF16 operands, one MFMA pass, FP32 accumulate). */
int fricc_contract(const char* term, const double* W, const int64_t W_shape[4], const int64_t W_strides[4],
                   fricc_sparse* T, double* out, const int64_t out_shape[4], const int64_t out_strides[4],
                   int precision, void* stream);

/* Same contraction with the dense operand already laid out as the kernel
 * reads it: A[r][x] row-major, r = the contracted index pair in T order,
 * x = W's free index pair in output order, element type per precision
 * (double / float / fricc_split_bf16 records / fricc_split_bf16_hi records).  Used by the device-resident update, which builds
 * its intermediates directly in this layout.  No counterpart in the reference. */
int fricc_contract_prepared(const char* term, const void* A, const int64_t W_shape[4], fricc_sparse* T,
                            double* out, const int64_t out_shape[4], const int64_t out_strides[4], int precision,
                            void* stream);

/* General form of the two calls above for any contraction of a dense 4-index
 * operand with T over two indices: spec "wwww,tttt->oooo" (one letter per
 * axis, exactly two letters shared by W and T), out += alpha * einsum(spec).
 * prepared != 0: W_or_A is the kernel layout A[r][x] (see above) in the
 * precision's element type; otherwise an FP64 array with W_strides.
 * out_precision: element type of out, FRICC_FP64 (double) or FRICC_FP32 (float).
 * Used for the t1-dressing of the ladder (cc_Wvvvv, fricc/friccsd.py:325-331)
 * applied to the sparse T instead of materialising W. */
int fricc_contract_spec(const char* spec, double alpha, const void* W_or_A, int prepared, const int64_t W_shape[4],
                        const int64_t W_strides[4], fricc_sparse* T, void* out, int out_precision,
                        const int64_t out_shape[4], const int64_t out_strides[4], int precision, void* stream);

/* C = alpha * einsum(spec, A, B) + beta * C for device tensors of any
 * strides (in elements; stride 0 broadcasts): spec "abc,cd->abd" style, one
 * letter per axis, at most 8 axes per tensor; letters in A, B and C are
 * batch, in A and C (B and C) rows (columns), in A and B summed.  Element
 * types *_precision FRICC_FP64 or FRICC_FP32 (converted to the compute type
 * on load); compute_precision FRICC_FP64 (FP64 MFMA) or FRICC_FP32 (FP32
 * MFMA).  Each group of letters must reduce to at most three strided levels
 * (FRICC_EINVAL otherwise).  beta == 0: C is not read.  Replaces the dense
 * NumPy/PySCF einsums of update_amps (fricc/friccsd.py:144-207, 309-376). */
int fricc_dense_contract(const char* spec, double alpha, const void* A, int a_precision, int a_ndim,
                         const int64_t* a_shape, const int64_t* a_strides, const void* B, int b_precision, int b_ndim,
                         const int64_t* b_shape, const int64_t* b_strides, double beta, void* C, int c_precision,
                         int c_ndim, const int64_t* c_shape, const int64_t* c_strides, int compute_precision,
                         void* stream);

/* C = alpha * einsum(spec, A, B) + beta * Csrc: fricc_dense_contract with the
 * beta term read from Csrc, a tensor of C's shape, strides and element type
 * (C itself, or one that does not overlap C), so an accumulation onto a
 * constant operand needs no copy of it into C first (the reference's
 * ``W = integrals + einsum(...)`` pattern, fricc/friccsd.py:194-207).  No
 * counterpart in the reference. */
int fricc_dense_contract_from(const char* spec, double alpha, const void* A, int a_precision, int a_ndim,
                              const int64_t* a_shape, const int64_t* a_strides, const void* B, int b_precision,
                              int b_ndim, const int64_t* b_shape, const int64_t* b_strides, double beta,
                              const void* Csrc, void* C, int c_precision, int c_ndim, const int64_t* c_shape,
                              const int64_t* c_strides, int compute_precision, void* stream);

/* (pq|rs) = sum_Q L[Q,p,q] L[Q,r,s] (sequential fma chain over Q) for
 * p in [rng[0],rng[1]), q in [rng[2],rng[3]), r in [rng[4],rng[5]),
 * s in [rng[6],rng[7]); written at out[(p-rng[0])*st[0] + (q-rng[2])*st[1] +
 * (r-rng[4])*st[2] + (s-rng[6])*st[3]].  L: device [naux][nmo][nmo].
 * Integral provider for synthetic density-fitted systems; the reference takes
 * PySCF integrals instead (fricc/friccsd.py:503-504). */
int fricc_eri_build(const double* L, int64_t naux, int64_t nmo, const int64_t rng[8], const int64_t strides[4],
                    double* out, void* stream);

/* Build (or reuse) the gathered view of T that a term or spec needs, so a
 * following contraction runs its GEMM only (timing, graph capture).  The view
 * is rebuilt automatically when T's entries change.  No counterpart in the
 * reference (which rescans T inside every kernel, src/sparse_tensor.cpp:66-68). */
int fricc_sparse_prepare(fricc_sparse* T, const char* term_or_spec, const int64_t W_shape[4],
                         const int64_t out_shape[4], int precision, void* stream);

/* BF16 operand for FRICC_BF16 contractions: A[R][X] row-major (FP64 when
 * src_precision == FRICC_FP64, FP32 when FRICC_FP32) -> out = 1 KB records
 * [ceil(X/256)][R][hi 256 | lo 256] bf16: element (r, x) has hi = bf16(a) at
 * ((x/256)*R + r)*512 + x%256 and lo = bf16(a - hi) 256 further; columns
 * past X are zero.  out holds R*ceil(X/256)*512 uint16.  No counterpart in
 * the reference (FP64 only). */
int fricc_split_bf16(const void* A, int src_precision, int64_t R, int64_t X, void* out, void* stream);

/* Operand of FRICC_BF16X1 contractions: A[R][X] (FP64 or FP32) -> 512-byte
 * records [ceil(X/256)][R][hi 256] bf16 (hi = bf16(a) at ((x/256)*R + r)*256
 * + x%256, columns past X zero); out holds R*ceil(X/256)*256 uint16.  No
 * counterpart in the reference. */
int fricc_split_bf16_hi(const void* A, int src_precision, int64_t R, int64_t X, void* out, void* stream);

/* fricc_split_bf16 of alpha*A + beta*B (A, B: same shape and precision; the
 * combination formed in that precision), so a combined operand needs no
 * temporary.  No counterpart in the reference (its update forms
 * 2*Wvoov - Wvovo implicitly by adding two contractions,
 * fricc/friccsd.py:247-251). */
int fricc_split_bf16_lincomb(const void* A, const void* B, int src_precision, double alpha, double beta, int64_t R,
                             int64_t X, void* out, void* stream);

/* fricc_split_bf16_lincomb into out and, from the same reads, fricc_split_bf16
 * of A into out_a and of B into out_b (either may be NULL): the three
 * operands one stage of the update reads (2 Wvoov - Wvovo, Wvoov, Wvovo),
 * with A and B read once.  No counterpart in the reference. */
int fricc_split_bf16_lincomb3(const void* A, const void* B, int src_precision, double alpha, double beta, int64_t R,
                              int64_t X, void* out, void* out_a, void* out_b, void* stream);

/* MFMA passes of FRICC_BF16 contractions: 3 = hi*hi + hi*lo + lo*hi (default,
 * about 16 significant bits per product), 1 = hi*hi only. */
int fricc_set_bf16_passes(int npass);

/* Fused end of the amplitude update: acc (in: direct T2 terms, out: t2new)
 *   t2new = (ovovT + acc + S + S^T(ij,ab->ji,ba) + t1[i,a] t1[j,b] wsum[a,b])
 *           / (eia[i,a] + eia[j,b])
 * and sums[0] = sum (t2new + t1new t1new) * lovov, sums[1] = sum over the
 * packed lower triangle of (t2new - t2old)^2 (device, FP64, block-ordered).
 * ovovT, acc, t2old, lovov: [o][o][v][v] FP64; S: same shape, FP64
 * (s_precision 0) or FP32 (1); t1, eia, t1new: [o][v]; wsum [v][v] or NULL.
 * lovov NULL means the exchange integral 2 ovovT[ijab] - ovovT[ijba] (the
 * reference's Lovov), formed in the kernel instead of read.  t2new_f32
 * [o][o][v][v] FP32 or NULL: also receives t2new rounded to FP32 (the
 * reduced-precision copy the next step's products read).
 * Replaces the symmetrisation, denominator, energy and norm passes of
 * fricc/friccsd.py:239-284 and :418-427. */
int fricc_t2_finalize(int o, int v, const double* ovovT, double* acc, const void* S, int s_precision,
                      const double* t1, const double* wsum, const double* eia, const double* t2old,
                      const double* lovov, const double* t1new, float* t2new_f32, double* sums, void* stream);

/* Two dense products that read the same o*v^3 operand Z[k][c][d][a], with Z
 * streamed once: out_ia[i][a] += sum_{k,c,d} t2[i][k][c][d] Z[k][c][d][a] and
 * out_ac[a][c] = sum_{k,d} t1[k][d] Z[k][c][d][a].  t2 [o][o][v][v], t1
 * [o][v], Z [o][v][v][v]: FP64 (precision FRICC_FP64, FP64 MFMA) or FP32
 * (FRICC_FP32, FP32 MFMA), C order, 16-byte aligned; out_ia [o][v] and
 * out_ac [v][v]: FP64.  Needs o <= 32 and v a multiple of 128 (FRICC_EINVAL
 * otherwise).  With Z = 2 (kd|ac) - (kc|ad) these are the t1 equation's
 * ovvv*t2 terms and the t1 part of Lvv (fricc/friccsd.py:165-168, 190),
 * which the reference computes as separate einsums; the fused entry point
 * has no counterpart in the reference. */
int fricc_dense_zpair(int o, int v, int precision, const void* t2, const void* t1, const void* Z, double* out_ia,
                      double* out_ac, void* stream);

/* Column-tile width of the gathered views of FRICC_BF16 contractions: 0 =
 * automatic (128), or a
 * multiple of 128.  Tuning knob; no counterpart in the reference. */
int fricc_set_bf16_view_width(int width);

/* Scratch arenas.  The library keeps grow-only scratch (compression buffers,
 * view slot tables, partial sums) per device in 64 numbered arenas; calls use
 * the arena the calling thread selected last (0 by default; the selection is
 * per host thread).  Work that can run concurrently on different streams --
 * solves served side by side -- must use different arenas; a recorded graph
 * keeps the addresses of the arena it was recorded with.
 * fricc_workspace_release frees an arena's buffers (it waits for the device).
 * fricc_workspace_pin(arena, +1) declares that a recorded graph holds the
 * arena, (arena, -1) that it was released: while an arena is pinned, a call
 * that would have to grow one of its slots fails with FRICC_EINVAL instead of
 * freeing memory the graph still uses, and release fails likewise.
 * No counterpart in the reference. */
int fricc_workspace_select(int arena);
int fricc_workspace_release(int arena);
int fricc_workspace_pin(int arena, int delta);

/* Device-to-device copy of `bytes` (a multiple of 16; 16-byte aligned device
 * pointers) by a plain streaming kernel: the reference point for achievable
 * HBM bandwidth and for scaling hardware traffic counters of the other
 * kernels.  No counterpart in the reference. */
int fricc_stream_copy(const void* src, void* dst, int64_t bytes, void* stream);

/* Block ordering of the gathered contraction GEMM: 0 = x tiles fastest,
 * 1 = column tiles fastest (default).  Tuning knob; no counterpart in the
 * reference. */
int fricc_set_gemm_order(int order);

/* K splits of the BF16 gathered contraction GEMM: 0 = automatic (split only
 * grids of fewer than 512 workgroups), or 1..16 splits of every tile's row
 * list, their partial sums added in split order.  Tuning knob; no
 * counterpart in the reference. */
int fricc_set_gemm_split(int nsplit);

#ifdef __cplusplus
}
#endif
#endif /* FRICC_CAPI_H */
