// Device algebra on CSR matrices: linear combinations (H(t) = H_static +
// sum_f f(t) H_f, hamiltonian_core.py:1686-1720 tocsr(time) and the
// hamiltonian __add__/__mul__ of _hamiltonian_core.py) and (conjugate)
// transposes (hamiltonian transpose/conj/getH, and the adjoint that
// expm_multiply_parallel's norm estimate needs), so that these stay in HBM
// instead of passing through host scipy.
//
// Inputs are canonical CSR: column indices sorted and unique within a row.
// Outputs are canonical too; a combination drops entries whose sum is exactly
// zero, as scipy's sparse addition does.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sparse_impl.h"

namespace {

constexpr int CSR_SUM_MAX = 32;

struct SumIn {
  const void* indptr[CSR_SUM_MAX];
  const void* indices[CSR_SUM_MAX];
  const void* data[CSR_SUM_MAX];
  int idx64[CSR_SUM_MAX];
  int dtype[CSR_SUM_MAX];
  int conj[CSR_SUM_MAX];
  double2 coef[CSR_SUM_MAX];
  int K;
};

__device__ __forceinline__ double2 ld_val(const void* p, int dt, int64_t k) {
  switch (dt) {
    case QSX_F32: return make_double2((double)((const float*)p)[k], 0.0);
    case QSX_F64: return make_double2(((const double*)p)[k], 0.0);
    case QSX_C64: { const float2 v = ((const float2*)p)[k]; return make_double2((double)v.x, (double)v.y); }
    default: return ((const double2*)p)[k];
  }
}

__device__ __forceinline__ void st_val(void* p, int dt, int64_t k, double2 v) {
  switch (dt) {
    case QSX_F32: ((float*)p)[k] = (float)v.x; break;
    case QSX_F64: ((double*)p)[k] = v.x; break;
    case QSX_C64: ((float2*)p)[k] = make_float2((float)v.x, (float)v.y); break;
    default: ((double2*)p)[k] = v; break;
  }
}

__device__ __forceinline__ void st_idx(void* p, int idx64, int64_t k, int64_t v) {
  if (idx64) ((int64_t*)p)[k] = v;
  else ((int32_t*)p)[k] = (int32_t)v;
}

// One row of sum_k coef_k A_k by a K-way merge of the sorted rows; FILL = 0
// counts the non-zero sums, FILL = 1 writes them from position out0.
template <int FILL>
__device__ void sum_row(const SumIn& S, int64_t r, int64_t& count, void* oix, void* oda, int oidx64, int odt,
                        int64_t out0) {
  int64_t pos[CSR_SUM_MAX], end[CSR_SUM_MAX];
  for (int k = 0; k < S.K; ++k) {
    pos[k] = ld_idx(S.indptr[k], S.idx64[k], r);
    end[k] = ld_idx(S.indptr[k], S.idx64[k], r + 1);
  }
  int64_t n = 0;
  for (;;) {
    int64_t c = INT64_MAX;
    for (int k = 0; k < S.K; ++k)
      if (pos[k] < end[k]) c = min(c, ld_idx(S.indices[k], S.idx64[k], pos[k]));
    if (c == INT64_MAX) break;
    double2 v = make_double2(0.0, 0.0);
    for (int k = 0; k < S.K; ++k) {
      if (pos[k] < end[k] && ld_idx(S.indices[k], S.idx64[k], pos[k]) == c) {
        double2 a = ld_val(S.data[k], S.dtype[k], pos[k]);
        if (S.conj[k]) a.y = -a.y;
        const double2 f = S.coef[k];
        v.x += f.x * a.x - f.y * a.y;
        v.y += f.x * a.y + f.y * a.x;
        ++pos[k];
      }
    }
    if (v.x != 0.0 || v.y != 0.0) {
      if constexpr (FILL) {
        st_idx(oix, oidx64, out0 + n, c);
        st_val(oda, odt, out0 + n, v);
      }
      ++n;
    }
  }
  count = n;
}

__global__ void __launch_bounds__(256) k_sum_count(SumIn S, int64_t Nr, int64_t* __restrict__ cnt) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < Nr; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t n;
    sum_row<0>(S, r, n, nullptr, nullptr, 0, 0, 0);
    cnt[r] = n;
  }
}

__global__ void __launch_bounds__(256) k_sum_fill(SumIn S, int64_t Nr, const int64_t* __restrict__ offs, void* oip,
                                                  void* oix, void* oda, int oidx64, int odt) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < Nr; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t n;
    sum_row<1>(S, r, n, oix, oda, oidx64, odt, offs[r]);
    st_idx(oip, oidx64, r, offs[r]);
    if (r == Nr - 1) st_idx(oip, oidx64, Nr, offs[r] + n);
  }
}

// transpose: sort keys col * Nr + row (stable, deterministic); values are the
// source entry positions
__global__ void __launch_bounds__(256) k_tr_keys(const void* indptr, const void* indices, int idx64, int64_t Nr,
                                                 uint64_t* __restrict__ keys, int64_t* __restrict__ vals) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < Nr; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k0 = ld_idx(indptr, idx64, r), k1 = ld_idx(indptr, idx64, r + 1);
    for (int64_t k = k0; k < k1; ++k) {
      keys[k] = (uint64_t)ld_idx(indices, idx64, k) * (uint64_t)Nr + (uint64_t)r;
      vals[k] = k;
    }
  }
}

__global__ void __launch_bounds__(256) k_tr_fill(const uint64_t* __restrict__ keys, const int64_t* __restrict__ vals,
                                                 int64_t nnz, int64_t Nr, int64_t Nc, const void* data, int dt,
                                                 int conj, void* oip, void* oix, void* oda, int oidx64, int odt) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[k];
    const int64_t c = (int64_t)(key / (uint64_t)Nr), r = (int64_t)(key % (uint64_t)Nr);
    double2 v = ld_val(data, dt, vals[k]);
    if (conj) v.y = -v.y;
    st_idx(oix, oidx64, k, r);
    st_val(oda, odt, k, v);
    // row starts of the transpose: the first entry of each column, and the
    // empty columns before it
    const int64_t cprev = k == 0 ? -1 : (int64_t)(keys[k - 1] / (uint64_t)Nr);
    for (int64_t j = cprev + 1; j <= c; ++j) st_idx(oip, oidx64, j, k);
    if (k == nnz - 1)
      for (int64_t j = c + 1; j <= Nc; ++j) st_idx(oip, oidx64, j, nnz);
  }
}

__global__ void k_tr_empty(void* oip, int oidx64, int64_t Nc) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= Nc; j += (int64_t)gridDim.x * blockDim.x)
    st_idx(oip, oidx64, j, 0);
}

int grid_for(int64_t n) {
  const int64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

bool dtype_ok(int d) { return d == QSX_F32 || d == QSX_F64 || d == QSX_C64 || d == QSX_C128; }

}  // namespace

struct qsx_csr_sum {
  SumIn S;
  int64_t Nr = 0;
  int64_t nnz = 0;
  int64_t* offs = nullptr;
};

extern "C" void qsx_csr_sum_destroy(qsx_csr_sum* H);

// sum_k coef_k op_k(A_k) of K canonical CSR matrices with Nr rows, op_k the
// identity or (conjs[k] = 1) the complex conjugate; coefs holds K (re, im)
// pairs.  Computes the row counts and returns nnz; qsx_csr_sum_finish writes
// the result.
extern "C" int qsx_csr_sum_create(int K, const void* const* indptrs, const void* const* indices,
                                  const void* const* datas, const int* idx64s, const int* dtypes, const int* conjs,
                                  const double* coefs, int64_t Nr, qsx_csr_sum** out, int64_t* nnz, void* stream) {
  if (!out || !nnz || K < 1 || K > CSR_SUM_MAX || Nr < 0 || Nr >= INT32_MAX || !indptrs || !indices || !datas ||
      !idx64s || !dtypes || !conjs || !coefs) {
    qsx_set_error("qsx_csr_sum_create: bad arguments");
    return QSX_ERR_ARG;
  }
  for (int k = 0; k < K; ++k)
    if (!dtype_ok(dtypes[k]) || !indptrs[k]) { qsx_set_error("qsx_csr_sum_create: bad matrix"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  auto* H = new qsx_csr_sum();
  std::memset(&H->S, 0, sizeof(SumIn));
  H->S.K = K;
  for (int k = 0; k < K; ++k) {
    H->S.indptr[k] = indptrs[k];
    H->S.indices[k] = indices[k];
    H->S.data[k] = datas[k];
    H->S.idx64[k] = idx64s[k];
    H->S.dtype[k] = dtypes[k];
    H->S.conj[k] = conjs[k] ? 1 : 0;
    H->S.coef[k] = make_double2(coefs[2 * k], coefs[2 * k + 1]);
  }
  H->Nr = Nr;
  *out = nullptr;
  *nnz = 0;
  if (Nr == 0) {
    *out = H;
    return QSX_OK;
  }
  // every failure below frees what was allocated and returns no handle
  int64_t* cnt = nullptr;
  void* tmp = nullptr;
  size_t tb = 0;
  int64_t lc = 0, lo = 0;
  hipError_t e = hipMalloc(&cnt, sizeof(int64_t) * Nr);
  if (e == hipSuccess) e = hipMalloc(&H->offs, sizeof(int64_t) * Nr);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sum_count, dim3(grid_for(Nr)), dim3(256), 0, st, H->S, Nr, cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, H->offs, (int)Nr, st);
  if (e == hipSuccess) e = hipMalloc(&tmp, tb + 16);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, H->offs, (int)Nr, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&lc, cnt + Nr - 1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&lo, H->offs + Nr - 1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(tmp);
  (void)hipFree(cnt);
  if (e != hipSuccess) {
    qsx_set_error(std::string("qsx_csr_sum_create: ") + hipGetErrorString(e));
    qsx_csr_sum_destroy(H);
    return QSX_ERR_HIP;
  }
  H->nnz = lc + lo;
  *out = H;
  *nnz = H->nnz;
  return QSX_OK;
}

// Writes the combination: out_indptr (Nr + 1), out_indices and out_data (nnz
// entries), indices of idx_bytes (4 or 8) bytes, values of out_dtype.
extern "C" int qsx_csr_sum_finish(qsx_csr_sum* H, int out_dtype, int idx_bytes, void* out_indptr, void* out_indices,
                                  void* out_data, void* stream) {
  if (!H || !dtype_ok(out_dtype) || (idx_bytes != 4 && idx_bytes != 8) || !out_indptr ||
      (H->nnz > 0 && (!out_indices || !out_data))) {
    qsx_set_error("qsx_csr_sum_finish: bad arguments");
    return QSX_ERR_ARG;
  }
  if (idx_bytes == 4 && (H->nnz >= INT32_MAX || H->Nr >= INT32_MAX)) {
    qsx_set_error("qsx_csr_sum_finish: 4-byte indices cannot hold this matrix");
    return QSX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int oidx64 = idx_bytes == 8;
  if (H->Nr == 0) {
    hipLaunchKernelGGL(k_tr_empty, dim3(1), dim3(256), 0, st, out_indptr, oidx64, (int64_t)0);
    QSX_LAUNCH_CHECK();
    return QSX_OK;
  }
  hipLaunchKernelGGL(k_sum_fill, dim3(grid_for(H->Nr)), dim3(256), 0, st, H->S, H->Nr, H->offs, out_indptr,
                     out_indices, out_data, oidx64, out_dtype);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" void qsx_csr_sum_destroy(qsx_csr_sum* H) {
  if (!H) return;
  if (H->offs) (void)hipFree(H->offs);
  delete H;
}

// (Conjugate) transpose of a canonical CSR with Nr rows, Nc columns and nnz
// entries into preallocated arrays: out_indptr (Nc + 1), out_indices and
// out_data (nnz), indices of idx_bytes bytes, values of out_dtype.
// Deterministic: entries of a row of the result keep ascending order.
extern "C" int qsx_csr_transpose(const void* indptr, const void* indices, const void* data, int idx64, int dtype,
                                 int64_t Nr, int64_t Nc, int64_t nnz, int conj, int out_dtype, int idx_bytes,
                                 void* out_indptr, void* out_indices, void* out_data, void* stream) {
  if (Nr < 0 || Nc < 0 || nnz < 0 || nnz >= INT32_MAX || !dtype_ok(dtype) || !dtype_ok(out_dtype) || (idx_bytes != 4 && idx_bytes != 8) ||
      !out_indptr || (nnz > 0 && (!indptr || !indices || !data || !out_indices || !out_data)) ||
      (Nr > 0 && Nc > 0 && (uint64_t)Nr > UINT64_MAX / (uint64_t)Nc)) {
    qsx_set_error("qsx_csr_transpose: bad arguments");
    return QSX_ERR_ARG;
  }
  if (idx_bytes == 4 && (nnz >= INT32_MAX || Nr >= INT32_MAX)) {
    qsx_set_error("qsx_csr_transpose: 4-byte indices cannot hold this matrix");
    return QSX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int oidx64 = idx_bytes == 8;
  if (nnz == 0) {
    hipLaunchKernelGGL(k_tr_empty, dim3(grid_for(Nc + 1)), dim3(256), 0, st, out_indptr, oidx64, Nc);
    QSX_LAUNCH_CHECK();
    return QSX_OK;
  }
  uint64_t *keys = nullptr, *keys2 = nullptr;
  int64_t *vals = nullptr, *vals2 = nullptr;
  void* tmp = nullptr;
  size_t tb = 0;
  int rc = QSX_OK;
  int end_bit = 1;
  const uint64_t kmax = (uint64_t)(Nc > 0 ? Nc : 1) * (uint64_t)(Nr > 0 ? Nr : 1);
  while (end_bit < 64 && (kmax >> end_bit) != 0) ++end_bit;
  hipError_t e = hipMalloc(&keys, sizeof(uint64_t) * nnz);
  if (e == hipSuccess) e = hipMalloc(&keys2, sizeof(uint64_t) * nnz);
  if (e == hipSuccess) e = hipMalloc(&vals, sizeof(int64_t) * nnz);
  if (e == hipSuccess) e = hipMalloc(&vals2, sizeof(int64_t) * nnz);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_tr_keys, dim3(grid_for(Nr)), dim3(256), 0, st, indptr, indices, idx64, Nr, keys, vals);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, keys2, vals, vals2, (int)nnz, 0,
                                                                end_bit, st);
  if (e == hipSuccess) e = hipMalloc(&tmp, tb + 16);
  if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, keys2, vals, vals2, (int)nnz, 0,
                                                                end_bit, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_tr_fill, dim3(grid_for(nnz)), dim3(256), 0, st, keys2, vals2, nnz, Nr, Nc, data, dtype,
                       conj, out_indptr, out_indices, out_data, oidx64, out_dtype);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    qsx_set_error(std::string("qsx_csr_transpose: ") + hipGetErrorString(e));
    rc = QSX_ERR_HIP;
  }
  (void)hipFree(tmp);
  (void)hipFree(keys);
  (void)hipFree(keys2);
  (void)hipFree(vals);
  (void)hipFree(vals2);
  return rc;
}
