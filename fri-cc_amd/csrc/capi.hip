// extern "C" entry points of libfricc_hip.so (declared in include/fricc_capi.h).
#include "../../include/fricc_capi.h"
#include "fricc_internal.h"

#include <string.h>

namespace fricc {
const char* last_error();
extern int g_order;
extern int g_split;
extern int g_bf16_width;

// The 11 contraction terms as index strings (src/sparse_tensor.cpp:11-346;
// the same identities the reference tests assert,
// fricc/tests/test_sparse_tensors.py:48-269).
struct TermSpec {
  const char* name;
  const char* w;
  const char* t;
  const char* o;
  double alpha;
};
static const TermSpec TERMS[] = {
    {"0101", "klij", "klab", "ijab", 1.0},  {"2323", "abcd", "ijcd", "ijab", 1.0},
    {"1302", "akic", "kjcb", "ijab", 2.0},  {"1202", "akci", "kjcb", "ijab", -1.0},
    {"1303", "akic", "kjbc", "ijab", 1.0},  {"1203", "bkci", "kjac", "ijab", 1.0},
    {"1323", "kcld", "ijcd", "klij", 1.0},  {"0112", "ldkc", "ilda", "akic", -0.5},
    {"0313", "lckd", "ilad", "akic", -0.5}, {"0113", "ldkc", "ilad", "akic", 1.0},
    {"0312", "lckd", "ilda", "akci", -0.5},
};

struct Plan {
  int pos[4];
  int64_t Rh, Rl, Xh, Xl;
  int wpos[4];  // W axes of rh, rl, xh, xl
  int opos[4];  // out axes of xh, xl, yh, yl
  double alpha;
};

static int find(const char* s, char c) {
  for (int i = 0; i < 4; i++)
    if (s[i] == c) return i;
  return -1;
}

static int make_plan_spec(const char* w, const char* t, const char* o, double alpha, const int64_t wshape[4],
                          const int64_t tdims[4], const int64_t oshape[4], Plan* p) {
  char r[4], y[4], x[4];
  int nr = 0, ny = 0, nx = 0;
  for (int i = 0; i < 4; i++) {
    char c = t[i];
    if (find(w, c) >= 0) {
      if (nr < 4) r[nr] = c;
      nr++;
    } else {
      if (ny < 4) y[ny] = c;
      ny++;
    }
  }
  for (int i = 0; i < 4; i++) {
    char c = o[i];
    if (find(w, c) >= 0) {
      if (nx < 4) x[nx] = c;
      nx++;
    }
  }
  FRICC_CHECK(nr == 2 && ny == 2 && nx == 2, "contract: spec must contract exactly two indices of W and T");
  FRICC_CHECK(find(o, y[0]) >= 0 && find(o, y[1]) >= 0, "contract: free T indices must appear in the output");
  p->pos[0] = find(t, r[0]);
  p->pos[1] = find(t, r[1]);
  p->pos[2] = find(t, y[0]);
  p->pos[3] = find(t, y[1]);
  p->wpos[0] = find(w, r[0]);
  p->wpos[1] = find(w, r[1]);
  p->wpos[2] = find(w, x[0]);
  p->wpos[3] = find(w, x[1]);
  p->opos[0] = find(o, x[0]);
  p->opos[1] = find(o, x[1]);
  p->opos[2] = find(o, y[0]);
  p->opos[3] = find(o, y[1]);
  p->Rh = tdims[p->pos[0]];
  p->Rl = tdims[p->pos[1]];
  p->Xh = wshape[p->wpos[2]];
  p->Xl = wshape[p->wpos[3]];
  p->alpha = alpha;
  FRICC_CHECK(wshape[p->wpos[0]] == p->Rh && wshape[p->wpos[1]] == p->Rl,
              "contract: W and T disagree on a contracted dimension");
  FRICC_CHECK(oshape[p->opos[0]] == p->Xh && oshape[p->opos[1]] == p->Xl,
              "contract: output and W disagree on a free dimension");
  FRICC_CHECK(oshape[p->opos[2]] == tdims[p->pos[2]] && oshape[p->opos[3]] == tdims[p->pos[3]],
              "contract: output and T disagree on a free dimension");
  return OK;
}

static int make_plan(const char* term, const int64_t wshape[4], const int64_t tdims[4], const int64_t oshape[4],
                     Plan* p) {
  const TermSpec* ts = nullptr;
  for (const auto& t : TERMS)
    if (term && strcmp(term, t.name) == 0) ts = &t;
  FRICC_CHECK(ts != nullptr, std::string("contract: term not implemented: ") + (term ? term : "(null)"));
  return make_plan_spec(ts->w, ts->t, ts->o, ts->alpha, wshape, tdims, oshape, p);
}

// "wxyz,tuvw->abcd" -> three 4-letter strings
static int parse_spec(const char* spec, char w[5], char t[5], char o[5]) {
  FRICC_CHECK(spec && strlen(spec) == 15 && spec[4] == ',' && spec[9] == '-' && spec[10] == '>',
              "contract: spec must look like 'abcd,efgh->ijkl'");
  memcpy(w, spec, 4);
  memcpy(t, spec + 5, 4);
  memcpy(o, spec + 11, 4);
  w[4] = t[4] = o[4] = 0;
  return OK;
}

static int dtype_of(int precision, int* dt) {
  if (precision == FRICC_FP64) *dt = DT_F64;
  else if (precision == FRICC_FP32) *dt = DT_F32;
  else if (precision == FRICC_BF16) *dt = DT_BF16;
  else if (precision == FRICC_BF16X1) *dt = DT_BF16X1;
  else FRICC_CHECK(false, "unknown precision");
  return OK;
}

static int g_npass = 3;

static int view_width(const Plan& p, int dt) {
  return (dt == DT_BF16 || dt == DT_BF16X1) ? bf16_view_width(p.Rh * p.Rl, p.Xh * p.Xl) : 64;
}

static int run_contract(const Plan& p, const void* A, int dt, Sparse* T, void* out, const int64_t os[4],
                        hipStream_t s, int of32 = 0) {
  View* v = nullptr;
  FRICC_TRY(get_view(T, p.pos, view_width(p, dt), dt, s, &v));
  OutMap om;
  om.out = out;
  om.of32 = of32;
  om.Xh = p.Xh;
  om.Xl = p.Xl;
  om.oxh = os[p.opos[0]];
  om.oxl = os[p.opos[1]];
  om.oyh = os[p.opos[2]];
  om.oyl = os[p.opos[3]];
  if (dt == DT_BF16) return gathered_gemm_bf16(v, (const uint16_t*)A, p.Rh * p.Rl, p.Xh * p.Xl, om, p.alpha, g_npass, s);
  if (dt == DT_BF16X1)
    return gathered_gemm_bf16(v, (const uint16_t*)A, p.Rh * p.Rl, p.Xh * p.Xl, om, p.alpha, 1, s, /*hilo=*/0);
  return gathered_gemm(v, A, dt, p.Xh * p.Xl, om, p.alpha, s);
}

// FP64/FP32 (or a permuted FP64 W) -> the BF16 operand (hi | lo records, or
// hi-only ones for DT_BF16X1) in workspace slot 31
static int bf16_operand_from_w(const Plan& p, const double* W, const int64_t W_strides[4], hipStream_t s,
                               const uint16_t** out, int dt) {
  const int64_t R = p.Rh * p.Rl, X = p.Xh * p.Xl;
  void* A = nullptr;
  void* A2 = nullptr;
  FRICC_TRY(ws_get(30, (size_t)R * X * 4, &A));
  FRICC_TRY(ws_get(31, (size_t)R * bf16_xblocks(X) * 512 * 2, &A2));
  const int64_t st[4] = {W_strides[p.wpos[0]], W_strides[p.wpos[1]], W_strides[p.wpos[2]], W_strides[p.wpos[3]]};
  FRICC_TRY(permute_operand(W, p.Rh, p.Rl, p.Xh, p.Xl, st, A, DT_F32, s));
  FRICC_TRY(split_bf16(A, DT_F32, R, X, (uint16_t*)A2, s, nullptr, 1.0, 0.0, dt == DT_BF16 ? 1 : 0));
  *out = (const uint16_t*)A2;
  return OK;
}

}  // namespace fricc

using namespace fricc;

struct fricc_sparse : public fricc::Sparse {};

extern "C" {

int fricc_abi_version(void) { return FRICC_ABI_VERSION; }
const char* fricc_last_error(void) { return fricc::last_error(); }

int fricc_compress(const double* x, int64_t n, int64_t m, int compression, double u, uint32_t* out_idx,
                   double* out_val, fricc_compress_info* info, void* stream) {
  CompressInfo ci;
  int st = compress(x, n, m, compression, u, out_idx, out_val, &ci, (hipStream_t)stream);
  if (info) memcpy(info, &ci, sizeof(ci));
  return st;
}

int fricc_compress_noalloc(const double* x, int64_t n, int64_t m, int compression, double u, uint32_t* out_idx,
                           double* out_val, fricc_compress_info* info, void* stream) {
  CompressInfo ci;
  int st = compress_noalloc(x, n, m, compression, u, out_idx, out_val, &ci, (hipStream_t)stream);
  if (info) memcpy(info, &ci, sizeof(ci));
  return st;
}

int64_t fricc_compress_noalloc_max_n(void) { return compress_noalloc_max_n(); }

int fricc_sample_systematic(const double* p, int64_t n, int64_t n_sample, double u, uint32_t* out_idx,
                            int64_t* n_notfound, void* stream) {
  return sample_systematic(p, n, n_sample, u, out_idx, n_notfound, (hipStream_t)stream);
}

int fricc_sample_pivotal(const double* p, int64_t n, int64_t n_sample, const double* uniforms, double u,
                         uint32_t* out_idx, void* stream) {
  return sample_pivotal(p, n, n_sample, uniforms, u, out_idx, (hipStream_t)stream);
}

int fricc_parallel_sample(const double* p, int64_t n, int64_t n_sample, int64_t n_chunks, const double* uniforms,
                          double u, uint32_t* out_idx, void* stream) {
  return parallel_sample(p, n, n_sample, n_chunks, uniforms, u, out_idx, (hipStream_t)stream);
}

int fricc_sparse_create(const int64_t dims[4], fricc_sparse** out) {
  FRICC_CHECK(out != nullptr, "sparse_create: null output");
  int64_t n = 1;
  for (int i = 0; i < 4; i++) {
    FRICC_CHECK(dims[i] > 0, "sparse_create: dimensions must be positive");
    n *= dims[i];
  }
  FRICC_CHECK(n < ((int64_t)1 << 31), "sparse_create: tensor too large (flat index must fit 31 bits)");
  fricc_sparse* sp = new fricc_sparse();
  for (int i = 0; i < 4; i++) sp->dims[i] = dims[i];
  FRICC_HIP(hipGetDevice(&sp->device));
  *out = sp;
  return OK;
}

int fricc_sparse_destroy(fricc_sparse* sp) {
  delete sp;
  return OK;
}

int fricc_sparse_compress(fricc_sparse* sp, const double* x, int64_t m, int compression, double u, void* stream) {
  FRICC_CHECK(sp != nullptr, "sparse_compress: null tensor");
  sp->pending = false;
  const int64_t n = sp->dims[0] * sp->dims[1] * sp->dims[2] * sp->dims[3];
  FRICC_TRY(sparse_reserve(sp, m));
  sp->gen++;
  sp->nnz = 0;
  sp->dup_dev = nullptr;
  sp->dup_host = 1;
  int dup = 1;
  int st = compress(x, n, m, compression, u, sp->idx, sp->val, &sp->info, (hipStream_t)stream, &dup);
  if (st == OK) {
    sp->nnz = m;
    sp->dup_host = dup;
  }
  return st;
}

int fricc_sparse_compress_async(fricc_sparse* sp, const double* x, int64_t m, int compression, const double* u_dev,
                                void* stream) {
  FRICC_CHECK(sp != nullptr && u_dev != nullptr, "sparse_compress_async: null argument");
  const int64_t n = sp->dims[0] * sp->dims[1] * sp->dims[2] * sp->dims[3];
  FRICC_TRY(sparse_reserve(sp, m));
  if (!sp->dsc) FRICC_HIP(hipMalloc(&sp->dsc, comp_scalars_size()));
  sp->gen++;
  sp->nnz = 0;
  void* sc = nullptr;
  FRICC_TRY(compress_launch(x, n, m, compression, 0.0, u_dev, sp->idx, sp->val, sp->dsc, (hipStream_t)stream, &sc));
  sp->nnz = m;
  sp->pending = true;
  sp->dup_host = 0;
  sp->dup_dev = comp_scalars_dup_flag(sp->dsc);
  return OK;
}

// Sample k >= 1 of one compression into sp2 (see compress_resample).
static int resample_into(fricc_sparse* first, fricc_sparse* sp2, const double* x, int64_t m, int compression,
                         const double* u_dev, hipStream_t s) {
  const int64_t n = sp2->dims[0] * sp2->dims[1] * sp2->dims[2] * sp2->dims[3];
  FRICC_TRY(sparse_reserve(sp2, m));
  if (!sp2->dsc) FRICC_HIP(hipMalloc(&sp2->dsc, comp_scalars_size()));
  sp2->gen++;
  sp2->nnz = 0;
  FRICC_TRY(compress_resample(x, n, m, compression, u_dev, first->idx, first->val, sp2->idx, sp2->val, sp2->dsc, s));
  sp2->nnz = m;
  sp2->pending = true;
  sp2->dup_host = 0;
  sp2->dup_dev = comp_scalars_dup_flag(sp2->dsc);
  return OK;
}

int fricc_sparse_compress_pair_async(fricc_sparse* sp, fricc_sparse* sp2, const double* x, int64_t m, int compression,
                                     const double* u_dev, const double* u_dev2, void* stream) {
  FRICC_CHECK(sp != nullptr && sp2 != nullptr && sp != sp2 && u_dev != nullptr && u_dev2 != nullptr,
              "sparse_compress_pair_async: two distinct tensors and two uniforms needed");
  for (int q = 0; q < 4; q++)
    FRICC_CHECK(sp->dims[q] == sp2->dims[q], "sparse_compress_pair_async: the tensors' dimensions differ");
  // one call for both samples: the selection once, both emissions in one pass
  const int64_t n = sp->dims[0] * sp->dims[1] * sp->dims[2] * sp->dims[3];
  for (fricc_sparse* t : {sp, sp2}) {
    FRICC_TRY(sparse_reserve(t, m));
    if (!t->dsc) FRICC_HIP(hipMalloc(&t->dsc, comp_scalars_size()));
    t->gen++;
    t->nnz = 0;
  }
  FRICC_TRY(compress_pair_launch(x, n, m, compression, u_dev, u_dev2, sp->idx, sp->val, sp->dsc, sp2->idx, sp2->val,
                                 sp2->dsc, (hipStream_t)stream));
  for (fricc_sparse* t : {sp, sp2}) {
    t->nnz = m;
    t->pending = true;
    t->dup_host = 0;
    t->dup_dev = comp_scalars_dup_flag(t->dsc);
  }
  return OK;
}

int fricc_sparse_compress_multi_async(fricc_sparse* const* sps, int count, const double* x, int64_t m,
                                      int compression, const double* u_dev, void* stream) {
  FRICC_CHECK(sps != nullptr && count >= 1 && u_dev != nullptr, "sparse_compress_multi_async: bad arguments");
  for (int k = 0; k < count; k++) {
    FRICC_CHECK(sps[k] != nullptr, "sparse_compress_multi_async: null tensor");
    for (int l = 0; l < k; l++) FRICC_CHECK(sps[l] != sps[k], "sparse_compress_multi_async: tensors must differ");
    for (int q = 0; q < 4; q++)
      FRICC_CHECK(sps[k]->dims[q] == sps[0]->dims[q], "sparse_compress_multi_async: the tensors' dimensions differ");
  }
  // samples 0 and 1 from one selection and one emission pass, then the
  // others two per emission pass (a last odd one alone)
  if (count == 1) return fricc_sparse_compress_async(sps[0], x, m, compression, u_dev, stream);
  FRICC_TRY(fricc_sparse_compress_pair_async(sps[0], sps[1], x, m, compression, u_dev, u_dev + 1, stream));
  const hipStream_t s = (hipStream_t)stream;
  const int64_t n = sps[0]->dims[0] * sps[0]->dims[1] * sps[0]->dims[2] * sps[0]->dims[3];
  int k = 2;
  for (; k + 1 < count; k += 2) {
    fricc_sparse* a = sps[k];
    fricc_sparse* b = sps[k + 1];
    for (fricc_sparse* t : {a, b}) {
      FRICC_TRY(sparse_reserve(t, m));
      if (!t->dsc) FRICC_HIP(hipMalloc(&t->dsc, comp_scalars_size()));
      t->gen++;
      t->nnz = 0;
    }
    FRICC_TRY(compress_resample_pair(x, n, m, compression, u_dev + k, u_dev + k + 1, sps[0]->idx, sps[0]->val, a->idx,
                                     a->val, a->dsc, b->idx, b->val, b->dsc, s));
    for (fricc_sparse* t : {a, b}) {
      t->nnz = m;
      t->pending = true;
      t->dup_host = 0;
      t->dup_dev = comp_scalars_dup_flag(t->dsc);
    }
  }
  if (k < count) FRICC_TRY(resample_into(sps[0], sps[k], x, m, compression, u_dev + k, s));
  return OK;
}

// Reads the scalars the last asynchronous compression left on the device
// (every replay of a recorded graph rewrites them, so they are read each time).
int fricc_sparse_status(fricc_sparse* sp, void* stream) {
  FRICC_CHECK(sp != nullptr, "sparse_status: null tensor");
  if (!sp->pending) return OK;
  std::vector<char> h(comp_scalars_size());
  hipStream_t s = (hipStream_t)stream;
  FRICC_HIP(hipMemcpyAsync(h.data(), sp->dsc, h.size(), hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  return comp_scalars_status(h.data(), &sp->info);
}

int fricc_sparse_set(fricc_sparse* sp, const int64_t* idx, const double* val, int64_t nnz, void* stream) {
  FRICC_CHECK(sp != nullptr && nnz >= 0, "sparse_set: bad arguments");
  const int64_t n = sp->dims[0] * sp->dims[1] * sp->dims[2] * sp->dims[3];
  std::vector<uint32_t> h(nnz);
  for (int64_t i = 0; i < nnz; i++) {
    FRICC_CHECK(idx[i] >= 0 && idx[i] < n, "sparse_set: index out of range (include/sparse_tensor.hpp:116-123)");
    h[i] = (uint32_t)idx[i];
  }
  FRICC_TRY(sparse_reserve(sp, nnz));
  sp->gen++;
  hipStream_t s = (hipStream_t)stream;
  if (nnz > 0) {
    FRICC_HIP(hipMemcpyAsync(sp->idx, h.data(), nnz * 4, hipMemcpyHostToDevice, s));
    FRICC_HIP(hipMemcpyAsync(sp->val, val, nnz * 8, hipMemcpyHostToDevice, s));
    FRICC_HIP(hipStreamSynchronize(s));
  }
  sp->nnz = nnz;
  sp->pending = false;
  sp->dup_host = 1;  // entries set by the caller may repeat an index
  sp->dup_dev = nullptr;
  memset(&sp->info, 0, sizeof(sp->info));
  sp->info.n_D = nnz;
  return OK;
}

int fricc_sparse_info(const fricc_sparse* sp, int64_t dims[4], int64_t* nnz, fricc_compress_info* info) {
  FRICC_CHECK(sp != nullptr, "sparse_info: null tensor");
  if (dims)
    for (int i = 0; i < 4; i++) dims[i] = sp->dims[i];
  if (nnz) *nnz = sp->nnz;
  if (info) memcpy(info, &sp->info, sizeof(*info));
  return OK;
}

int fricc_sparse_get(const fricc_sparse* sp, int64_t* idx, double* val, void* stream) {
  FRICC_CHECK(sp != nullptr, "sparse_get: null tensor");
  const int64_t nnz = sp->nnz;
  if (nnz == 0) return OK;
  std::vector<uint32_t> h(nnz);
  hipStream_t s = (hipStream_t)stream;
  FRICC_HIP(hipMemcpyAsync(h.data(), sp->idx, nnz * 4, hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipMemcpyAsync(val, sp->val, nnz * 8, hipMemcpyDeviceToHost, s));
  FRICC_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < nnz; i++) idx[i] = h[i];
  return OK;
}

int fricc_sparse_device(const fricc_sparse* sp, const uint32_t** idx, const double** val) {
  FRICC_CHECK(sp != nullptr, "sparse_device: null tensor");
  if (idx) *idx = sp->idx;
  if (val) *val = sp->val;
  return OK;
}

int fricc_contract(const char* term, const double* W, const int64_t W_shape[4], const int64_t W_strides[4],
                   fricc_sparse* T, double* out, const int64_t out_shape[4], const int64_t out_strides[4],
                   int precision, void* stream) {
  FRICC_CHECK(T != nullptr && W != nullptr && out != nullptr, "contract: null argument");
  Plan p;
  FRICC_TRY(make_plan(term, W_shape, T->dims, out_shape, &p));
  int dt;
  FRICC_TRY(dtype_of(precision, &dt));
  hipStream_t s = (hipStream_t)stream;
  if (dt == DT_BF16 || dt == DT_BF16X1) {
    const uint16_t* A2 = nullptr;
    FRICC_TRY(bf16_operand_from_w(p, W, W_strides, s, &A2, dt));
    return run_contract(p, A2, dt, T, out, out_strides, s);
  }
  const int64_t R = p.Rh * p.Rl, X = p.Xh * p.Xl;
  void* A = nullptr;
  FRICC_TRY(ws_get(30, (size_t)R * X * (dt == DT_F64 ? 8 : 4), &A));
  const int64_t st[4] = {W_strides[p.wpos[0]], W_strides[p.wpos[1]], W_strides[p.wpos[2]], W_strides[p.wpos[3]]};
  FRICC_TRY(permute_operand(W, p.Rh, p.Rl, p.Xh, p.Xl, st, A, dt, s));
  return run_contract(p, A, dt, T, out, out_strides, s);
}

int fricc_contract_prepared(const char* term, const void* A, const int64_t W_shape[4], fricc_sparse* T, double* out,
                            const int64_t out_shape[4], const int64_t out_strides[4], int precision, void* stream) {
  FRICC_CHECK(T != nullptr && A != nullptr && out != nullptr, "contract_prepared: null argument");
  Plan p;
  FRICC_TRY(make_plan(term, W_shape, T->dims, out_shape, &p));
  int dt;
  FRICC_TRY(dtype_of(precision, &dt));
  return run_contract(p, A, dt, T, out, out_strides, (hipStream_t)stream);
}

int fricc_contract_spec(const char* spec, double alpha, const void* W_or_A, int prepared, const int64_t W_shape[4],
                        const int64_t W_strides[4], fricc_sparse* T, void* out, int out_precision,
                        const int64_t out_shape[4], const int64_t out_strides[4], int precision, void* stream) {
  FRICC_CHECK(T != nullptr && W_or_A != nullptr && out != nullptr, "contract_spec: null argument");
  char w[5], t[5], o[5];
  FRICC_TRY(parse_spec(spec, w, t, o));
  Plan p;
  FRICC_TRY(make_plan_spec(w, t, o, alpha, W_shape, T->dims, out_shape, &p));
  int dt;
  FRICC_TRY(dtype_of(precision, &dt));
  FRICC_CHECK(out_precision == FRICC_FP64 || out_precision == FRICC_FP32, "contract_spec: output must be FP64 or FP32");
  const int of32 = out_precision == FRICC_FP32;
  hipStream_t s = (hipStream_t)stream;
  if (prepared) return run_contract(p, W_or_A, dt, T, out, out_strides, s, of32);
  if (dt == DT_BF16 || dt == DT_BF16X1) {
    const uint16_t* A2 = nullptr;
    FRICC_TRY(bf16_operand_from_w(p, (const double*)W_or_A, W_strides, s, &A2, dt));
    return run_contract(p, A2, dt, T, out, out_strides, s, of32);
  }
  const int64_t R = p.Rh * p.Rl, X = p.Xh * p.Xl;
  void* A = nullptr;
  FRICC_TRY(ws_get(30, (size_t)R * X * (dt == DT_F64 ? 8 : 4), &A));
  const int64_t st[4] = {W_strides[p.wpos[0]], W_strides[p.wpos[1]], W_strides[p.wpos[2]], W_strides[p.wpos[3]]};
  FRICC_TRY(permute_operand((const double*)W_or_A, p.Rh, p.Rl, p.Xh, p.Xl, st, A, dt, s));
  return run_contract(p, A, dt, T, out, out_strides, s, of32);
}

int fricc_sparse_prepare(fricc_sparse* T, const char* spec, const int64_t W_shape[4], const int64_t out_shape[4],
                         int precision, void* stream) {
  FRICC_CHECK(T != nullptr, "sparse_prepare: null tensor");
  Plan p;
  if (spec && strlen(spec) == 4) {
    FRICC_TRY(make_plan(spec, W_shape, T->dims, out_shape, &p));
  } else {
    char w[5], t[5], o[5];
    FRICC_TRY(parse_spec(spec, w, t, o));
    FRICC_TRY(make_plan_spec(w, t, o, 1.0, W_shape, T->dims, out_shape, &p));
  }
  int dt;
  FRICC_TRY(dtype_of(precision, &dt));
  View* v = nullptr;
  return get_view(T, p.pos, view_width(p, dt), dt, (hipStream_t)stream, &v);
}

int fricc_split_bf16(const void* A, int src_precision, int64_t R, int64_t X, void* out, void* stream) {
  int dt;
  FRICC_TRY(dtype_of(src_precision, &dt));
  return split_bf16(A, dt, R, X, (uint16_t*)out, (hipStream_t)stream);
}

int fricc_split_bf16_hi(const void* A, int src_precision, int64_t R, int64_t X, void* out, void* stream) {
  int dt;
  FRICC_TRY(dtype_of(src_precision, &dt));
  return split_bf16(A, dt, R, X, (uint16_t*)out, (hipStream_t)stream, nullptr, 1.0, 0.0, 0);
}

int fricc_split_bf16_lincomb(const void* A, const void* B, int src_precision, double alpha, double beta, int64_t R,
                             int64_t X, void* out, void* stream) {
  FRICC_CHECK(A != nullptr && B != nullptr, "split_bf16_lincomb: null operand");
  int dt;
  FRICC_TRY(dtype_of(src_precision, &dt));
  return split_bf16(A, dt, R, X, (uint16_t*)out, (hipStream_t)stream, B, alpha, beta);
}

int fricc_split_bf16_lincomb3(const void* A, const void* B, int src_precision, double alpha, double beta, int64_t R,
                              int64_t X, void* out, void* out_a, void* out_b, void* stream) {
  FRICC_CHECK(A != nullptr && B != nullptr && out != nullptr, "split_bf16_lincomb3: null operand");
  int dt;
  FRICC_TRY(dtype_of(src_precision, &dt));
  return split_bf16(A, dt, R, X, (uint16_t*)out, (hipStream_t)stream, B, alpha, beta, 1, (uint16_t*)out_a,
                    (uint16_t*)out_b);
}

int fricc_set_bf16_passes(int npass) {
  FRICC_CHECK(npass == 1 || npass == 3, "set_bf16_passes: 1 or 3");
  g_npass = npass;
  return OK;
}

// Block ordering of the gathered GEMM (0: x tiles fastest, 1: column tiles
// fastest).  Tuning knob; no counterpart in the reference.
int fricc_set_bf16_view_width(int width) {
  FRICC_CHECK(width == 0 || (width > 0 && width % 128 == 0 && width <= 4096),
              "set_bf16_view_width: 0 (automatic) or a multiple of 128 up to 4096");
  g_bf16_width = width;
  return OK;
}

int fricc_set_gemm_order(int order) {
  FRICC_CHECK(order == 0 || order == 1, "set_gemm_order: order must be 0 or 1");
  g_order = order;
  return OK;
}

int fricc_set_gemm_split(int nsplit) {
  FRICC_CHECK(nsplit >= 0 && nsplit <= 16, "set_gemm_split: 0 (automatic) to 16");
  g_split = nsplit;
  return OK;
}

}  // extern "C"
