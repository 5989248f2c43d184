// Internal (C++) interfaces shared by the HIP translation units.
#pragma once
#include "common.h"

#include <map>
#include <memory>
#include <tuple>
#include <vector>

namespace fricc {

enum { COMP_LARGEST = 0, COMP_FRI_SYSTEMATIC = 1, COMP_FRI_PIVOTAL = 2 };
// DT_BF16: hi | lo bf16 planes (three-pass products); DT_BF16X1: the hi plane
// only (single-pass products)
enum { DT_F64 = 0, DT_F32 = 1, DT_BF16 = 2, DT_BF16X1 = 3 };

struct CompressInfo {
  int64_t n_D, n_sampled, n_notfound, n_candidates;
  double norm0, remaining_norm, p_total;
  int32_t nonfinite, pad;
};

// dup (optional): set when the output may repeat an index
int compress(const double* x, int64_t n, int64_t m, int method, double u, uint32_t* out_idx,
             double* out_val, CompressInfo* info, hipStream_t s, int* dup = nullptr);
// The same compression with no host round trip (u from u_dev when given);
// the call's scalars go to export_sc (device, comp_scalars_size() bytes).
int compress_launch(const double* x, int64_t n, int64_t m, int method, double u_val, const double* u_dev,
                    uint32_t* out_idx, double* out_val, void* export_sc, hipStream_t s, void** sc_out);
// A second sample, with the uniform u_dev, of the vector compress_launch has
// just compressed into idx/val (same x, n, m, method, stream and arena).
int compress_resample(const double* x, int64_t n, int64_t m, int method, const double* u_dev, const uint32_t* idx,
                      const double* val, uint32_t* out_idx, double* out_val, void* export_sc, hipStream_t s);
// Two samples (u_dev, u_dev2) of one compression in one call: bit for bit
// compress_launch(u_dev) then compress_resample(u_dev2).
int compress_pair_launch(const double* x, int64_t n, int64_t m, int method, const double* u_dev,
                         const double* u_dev2, uint32_t* idx1, double* val1, void* export1, uint32_t* idx2,
                         double* val2, void* export2, hipStream_t s);
// Two more samples (u_a, u_b) after compress_launch / compress_pair_launch:
// bit for bit two compress_resample calls.
int compress_resample_pair(const double* x, int64_t n, int64_t m, int method, const double* u_a, const double* u_b,
                           const uint32_t* idx, const double* val, uint32_t* idx_a, double* val_a, void* export_a,
                           uint32_t* idx_b, double* val_b, void* export_b, hipStream_t s);
// The same as compress(), with all scratch in a fixed device array (no
// allocation); n <= compress_noalloc_max_n().
int compress_noalloc(const double* x, int64_t n, int64_t m, int method, double u, uint32_t* out_idx,
                     double* out_val, CompressInfo* info, hipStream_t s);
int64_t compress_noalloc_max_n();
size_t comp_scalars_size();
// device address of the "output may repeat an index" word of a call's scalars
const uint32_t* comp_scalars_dup_flag(const void* dev_sc);
// statistics of a call's scalars (host copy) into info; returns the call's status
int comp_scalars_status(const void* host_sc, CompressInfo* info);
int sample_systematic(const double* p, int64_t n, int64_t n_sample, double u, uint32_t* out_idx,
                      int64_t* n_notfound, hipStream_t s);
// Pivotal sampling of n_sample indices from p[n]; step j draws U[j] (device,
// n doubles) or, with U null, SplitMix64 keyed by the bits of u.
int sample_pivotal(const double* p, int64_t n, int64_t n_sample, const double* U, double u, uint32_t* out_idx,
                   hipStream_t s);
// parallel_sample with T chunks (draws by position as in sample_pivotal).
int parallel_sample(const double* q, int64_t n, int64_t n_sample, int64_t T, const double* U, double u,
                    uint32_t* out_idx, hipStream_t s);

// A gathered view of a sparse 4-index tensor as a matrix Tm[r, y]:
// rows r = (i_rh, i_rl), columns y = (i_yh, i_yl) (positions into the tensor's
// four indices).  Columns are cut into tiles of TY; for each tile the rows
// that hold any entry are gathered and stored densely ([rows x TY]).
struct View {
  int pos[4];          // rh, rl, yh, yl
  int64_t Rh, Rl, Yh, Yl, R, Y;
  int TY, kstep, dtype;
  int64_t ntiles = 0, nnz = 0;
  uint32_t* tile_off = nullptr;  // [ntiles + 1] first gathered row of each tile (device; last = total)
  uint32_t* tile_order = nullptr;  // [ntiles] tiles by decreasing row count (device)
  uint32_t* rowlist = nullptr;   // [cap_rows] r of each gathered row
  void* B = nullptr;             // [rows x TY] values (dtype); BF16: hi plane, bf16 [rows/8][TY][8]
  void* Blo = nullptr;           // DT_BF16 only: lo plane (value - hi), same layout
  size_t cap_rows = 0;           // worst-case rows the buffers hold
  uint64_t gen = 0;
  ~View();
};

struct Sparse {
  int64_t dims[4] = {0, 0, 0, 0};
  int64_t nnz = 0, cap = 0;
  uint32_t* idx = nullptr;   // flat index ((i*d1+j)*d2+a)*d3+b
  double* val = nullptr;
  CompressInfo info{};
  void* dsc = nullptr;   // device copy of the last asynchronous compression's scalars
  bool pending = false;  // info/status of an asynchronous compression not read yet
  // whether the entries may repeat a flat index (then views add them up):
  // dup_host, or the device word dup_dev when set (asynchronous compressions)
  int dup_host = 1;
  const uint32_t* dup_dev = nullptr;
  int device = 0;
  uint64_t gen = 1;  // bumped whenever the entries change
  std::map<std::tuple<int, int, int, int, int, int>, std::unique_ptr<View>> views;
  ~Sparse();
};

int sparse_reserve(Sparse* sp, int64_t nnz);
int get_view(Sparse* sp, const int pos[4], int TY, int dtype, hipStream_t s, View** out);

// OUT[x, y] += alpha * sum_k A[rowlist[k], x] * B[k, y]   (per column tile)
// A: row-major [R x X] in `adtype` (F64 or F32 here; BF16 uses its own path).
// Output element (x, y) lives at out + (x/Xl)*oxh + (x%Xl)*oxl + (y/Yl)*oyh + (y%Yl)*oyl.
struct OutMap {
  void* out;      // double*, or float* when of32
  int of32 = 0;   // output element type: 0 FP64, 1 FP32
  int64_t Xh, Xl, oxh, oxl, oyh, oyl;
};
int gathered_gemm(const View* v, const void* A, int adtype, int64_t X, const OutMap& om, double alpha,
                  hipStream_t s);

// BF16 views gather single rows; their planes interleave each 8 consecutive
// gathered rows per column.  A is given as 1 KB records
// [bf16_xblocks(X)][R][hi 256 | lo 256] bf16.  npass 3 = hi*hi + hi*lo +
// lo*hi, 1 = hi*hi.
// hilo 0: A holds hi-only records [bf16_xblocks(X)][R][hi 256] (npass 1)
int gathered_gemm_bf16(const View* v, const uint16_t* A2, int64_t R, int64_t X, const OutMap& om, double alpha,
                       int npass, hipStream_t s, int hilo = 1);
// tile width of BF16 views for an operand of R x X
int bf16_view_width(int64_t R, int64_t X);
// FP64/FP32 row-major A[R][X] -> [hi | lo] bf16 planes (layout above)
// B (optional): split alpha*A + beta*B instead of A; then A2a / A2b
// (optional): the records of A and of B as well, from the same reads
// hilo 0: hi-only records [bf16_xblocks(X)][R][hi 256]
int split_bf16(const void* A, int src_dtype, int64_t R, int64_t X, uint16_t* A2, hipStream_t s,
               const void* B = nullptr, double alpha = 1.0, double beta = 0.0, int hilo = 1, uint16_t* A2a = nullptr,
               uint16_t* A2b = nullptr);
// 256-column blocks of a BF16 operand (its size is R * blocks * 512 bf16)
inline int64_t bf16_xblocks(int64_t X) { return (X + 255) / 256; }

// A[r, x] = W[rh*s0 + rl*s1 + xh*s2 + xl*s3] (r = rh*Rl + rl, x = xh*Xl + xl) in dtype.
int permute_operand(const double* W, int64_t Rh, int64_t Rl, int64_t Xh, int64_t Xl, const int64_t s[4],
                    void* A, int dtype, hipStream_t s_);

}  // namespace fricc
