// CSR H.v, trace / 1-norm, device-resident Taylor-series expm and the
// FP64-accumulating vector kernels used by Lanczos.
//
// Replaces parallel_sparse_tools ``_matvec`` (hamiltonian_core.py:607-612,
// :741-772), ``_wrapper_csr_trace`` / ``_wrapper_csr_1_norm`` (EMP.html:196-202)
// and ``_wrapper_expm_multiply[_batch]`` (EMP.html:347-372), whose loop is
// Al-Mohy & Higham Alg. 3.2 as in scipy ``_expm_multiply_simple_core``
// (scipy/sparse/linalg/_expm_multiply.py:280-303).
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "sparse_impl.h"

// ---------------------------------------------------------------------------
// CSR SpMV dispatch
// ---------------------------------------------------------------------------
// Max over the per-block partials of one expm SpMV launch into the norm
// slots (a few hundred atomics per launch instead of one per SpMV block).
__global__ void __launch_bounds__(256) k_part_reduce(const double* __restrict__ part, int64_t nblocks,
                                                     const int* done, double* normB, double* normF, int nsp) {
  if (*done) return;
  double mb = 0.0, mf = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nblocks; i += (int64_t)gridDim.x * blockDim.x) {
    mb = fmax(mb, part[2 * i]);
    mf = fmax(mf, part[2 * i + 1]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mb = fmax(mb, __shfl_xor(mb, o, 64));
    mf = fmax(mf, __shfl_xor(mf, o, 64));
  }
  __shared__ double sw[2][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sw[0][w] = mb; sw[1][w] = mf; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) { mb = fmax(mb, sw[0][i]); mf = fmax(mf, sw[1][i]); }
    slot_put(normB, nsp, mb);
    slot_put(normF, nsp, mf);
  }
}

// Fold the per-block norm partials of one expm launch (nblocks blocks) into
// its norm slots; used by every expm term kernel.
int qsx_part_reduce_launch(const CsrArgs& A, int64_t nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_part_reduce, dim3(64), dim3(256), 0, st, A.part, nblocks, &A.ctl->done, A.normB, A.normF,
                     A.nsp);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// SpMV launches stride over rows with at most this many blocks (measured
// faster than one block per 256/G rows on 10^7-row matrices; the per-block
// norm partials of the expm form then fit a fixed buffer)
constexpr int64_t EPI_MAX_BLOCKS = 65536;

static int64_t csr_blocks(int64_t Ns, int G, bool /*epi*/) {
  int64_t nb = qsx_grid(Ns * G, 256);
  if (nb > EPI_MAX_BLOCKS) nb = EPI_MAX_BLOCKS;
  return nb < 1 ? 1 : nb;
}

template <typename MT, typename VT, bool EPI, typename FT = VT>
static void launch_csr_g(const CsrArgs& A, int G, hipStream_t st) {
  const dim3 grid((unsigned)csr_blocks(A.Ns, G, EPI)), blk(256);
  switch (G) {
    case 1: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 1, EPI, FT>), grid, blk, 0, st, A); break;
    case 2: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 2, EPI, FT>), grid, blk, 0, st, A); break;
    case 4: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 4, EPI, FT>), grid, blk, 0, st, A); break;
    case 8: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 8, EPI, FT>), grid, blk, 0, st, A); break;
    case 16: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 16, EPI, FT>), grid, blk, 0, st, A); break;
    case 32: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 32, EPI, FT>), grid, blk, 0, st, A); break;
    default: hipLaunchKernelGGL((k_csr_spmv<MT, VT, 64, EPI, FT>), grid, blk, 0, st, A); break;
  }
}

// fdt: dtype of the expm accumulator F (-1: the vector dtype); FP64 F with
// FP32 vectors is the mixed-precision form.
template <bool EPI>
static int launch_csr(const CsrArgs& A, int mdt, int vdt, int G, hipStream_t st, int fdt = -1) {
  if (fdt < 0) fdt = vdt;
  // matrix dtype must be the vector dtype or its real counterpart
  if (fdt == vdt) {
    if (vdt == QSX_F32 && mdt == QSX_F32) launch_csr_g<float, float, EPI>(A, G, st);
    else if (vdt == QSX_F64 && mdt == QSX_F64) launch_csr_g<double, double, EPI>(A, G, st);
    else if (vdt == QSX_C64 && mdt == QSX_F32) launch_csr_g<float, float2, EPI>(A, G, st);
    else if (vdt == QSX_C64 && mdt == QSX_C64) launch_csr_g<float2, float2, EPI>(A, G, st);
    else if (vdt == QSX_C128 && mdt == QSX_F64) launch_csr_g<double, double2, EPI>(A, G, st);
    else if (vdt == QSX_C128 && mdt == QSX_C128) launch_csr_g<double2, double2, EPI>(A, G, st);
    else { qsx_set_error("unsupported (matrix, vector) dtype pair"); return QSX_ERR_ARG; }
  } else if constexpr (EPI) {
    if (vdt == QSX_F32 && fdt == QSX_F64 && mdt == QSX_F32) launch_csr_g<float, float, EPI, double>(A, G, st);
    else if (vdt == QSX_C64 && fdt == QSX_C128 && mdt == QSX_F32) launch_csr_g<float, float2, EPI, double2>(A, G, st);
    else if (vdt == QSX_C64 && fdt == QSX_C128 && mdt == QSX_C64) launch_csr_g<float2, float2, EPI, double2>(A, G, st);
    else { qsx_set_error("unsupported mixed-precision (matrix, terms, accumulator) dtypes"); return QSX_ERR_ARG; }
  } else {
    qsx_set_error("accumulator dtype differs from the vector dtype outside the expm");
    return QSX_ERR_ARG;
  }
  QSX_LAUNCH_CHECK();
  if constexpr (EPI) return qsx_part_reduce_launch(A, csr_blocks(A.Ns, G, true), st);
  return QSX_OK;
}

// Lanes per row from the mean row length: about a third to a sixth of it, at
// least 2 (measured on MI355X: 16 entries per row run best with 4 lanes).
static int auto_group(int64_t Ns, int64_t nnz) {
  const double avg = Ns > 0 ? (double)nnz / (double)Ns : 1.0;
  if (avg <= 2.0) return 1;
  int G = 2;
  while (G < 64 && 3.0 * (2 * G) <= avg) G <<= 1;
  return G;
}

// y = a*A*x (+ y when accumulate) ; x, y: (Ns, nvec) row-major.
extern "C" int qsx_csr_matvec(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                              int64_t Ns, int64_t nnz, const void* x, void* y, int vec_dtype, int64_t nvec,
                              double a_re, double a_im, int accumulate, int group, void* stream) {
  if (Ns < 0 || nvec < 1) { qsx_set_error("qsx_csr_matvec: bad arguments"); return QSX_ERR_ARG; }
  if (Ns == 0) return QSX_OK;
  if ((vec_dtype == QSX_F32 || vec_dtype == QSX_F64) && a_im != 0.0) {
    qsx_set_error("real vectors with a complex scale"); return QSX_ERR_COMPLEX;
  }
  CsrArgs A;
  std::memset(&A, 0, sizeof(A));
  A.indptr = indptr; A.indices = indices; A.data = data; A.idx64 = idx64; A.Ns = Ns; A.nvec = nvec;
  A.x = x; A.y = y; A.a = make_double2(a_re, a_im); A.mu = make_double2(0, 0); A.beta = accumulate ? 1 : 0;
  const int G = group > 0 ? group : auto_group(Ns, nnz);
  return launch_csr<false>(A, mat_dtype, vec_dtype, G, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// deterministic block reductions (last block sums the partials in order)
// ---------------------------------------------------------------------------
template <int OP>  // 0: sum into out[0..1], 1: max into out[0]
__device__ __forceinline__ void block_finish(double vr, double vi, double* partials, unsigned int* counter,
                                             double* out, double post_scale_sqrt) {
  __shared__ double sr[256], si[256];
  __shared__ bool last;
  sr[threadIdx.x] = vr;
  si[threadIdx.x] = vi;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      if (OP == 0) { sr[threadIdx.x] += sr[threadIdx.x + o]; si[threadIdx.x] += si[threadIdx.x + o]; }
      else sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sr[0];
    partials[2 * blockIdx.x + 1] = si[0];
    __threadfence();
    const unsigned int t = atomicAdd(counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last) {
    // the last block reduces every block's partial: threads read a fixed
    // strided share (device-coherent loads, past this CU's cache), then the
    // same fixed tree as above, so the result does not depend on timing
    __threadfence();
    double a = 0.0, b = 0.0;
    for (unsigned int k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
      const double pr = __hip_atomic_load(partials + 2 * k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const double pi = __hip_atomic_load(partials + 2 * k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (OP == 0) { a += pr; b += pi; } else a = fmax(a, pr);
    }
    sr[threadIdx.x] = a;
    si[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        if (OP == 0) { sr[threadIdx.x] += sr[threadIdx.x + o]; si[threadIdx.x] += si[threadIdx.x + o]; }
        else sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      a = sr[0];
      if (post_scale_sqrt != 0.0) a = sqrt(a);
      out[0] = a;
      out[1] = si[0];
      *counter = 0u;
    }
  }
}

constexpr int RED_BLOCKS = 512;

// out = sum conj(x) y  (complex: out[0]+i out[1])
template <typename VT>
__global__ void __launch_bounds__(256) k_vdot(const VT* __restrict__ x, const VT* __restrict__ y, int64_t n,
                                              double* partials, unsigned int* counter, double* out) {
  double ar = 0.0, ai = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = VTr<VT>::ld(x + i), b = VTr<VT>::ld(y + i);
    ar += a.x * b.x + a.y * b.y;
    ai += a.x * b.y - a.y * b.x;
  }
  block_finish<0>(ar, ai, partials, counter, out, 0.0);
}

// y += a x with a = sgn * s[0] (+ i s[1]) read from device memory (or host value)
template <typename VT>
__global__ void __launch_bounds__(256) k_axpy_dev(const VT* __restrict__ x, VT* __restrict__ y, int64_t n,
                                                  const double* __restrict__ sdev, double sgn, double2 ahost,
                                                  int cplx_scale) {
  double2 a = ahost;
  if (sdev) a = make_double2(sgn * sdev[0], cplx_scale ? sgn * sdev[1] : 0.0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = VTr<VT>::ld(y + i);
    const double2 u = VTr<VT>::ld(x + i);
    if constexpr (VTr<VT>::cplx) v = cadd(v, cmul(a, u));
    else v.x += a.x * u.x;
    VTr<VT>::st(y + i, v);
  }
}

// Lanczos: r -= b_prev * v (b_prev device scalar, may be null);
// a = Re <q, r>  -> aout[0]
template <typename VT>
__global__ void __launch_bounds__(256) k_lz_axpy_dot(const VT* __restrict__ v, const double* __restrict__ bprev,
                                                     const VT* __restrict__ q, VT* __restrict__ r, int64_t n,
                                                     double* partials, unsigned int* counter, double* aout) {
  const double b = bprev ? bprev[0] : 0.0;
  double ar = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 rv = VTr<VT>::ld(r + i);
    if (bprev) {
      const double2 vv = VTr<VT>::ld(v + i);
      rv.x -= b * vv.x;
      rv.y -= b * vv.y;
      VTr<VT>::st(r + i, rv);
      rv = VTr<VT>::rnd(rv);  // the stored value, without a reload
    }
    const double2 qq = VTr<VT>::ld(q + i);
    ar += qq.x * rv.x + qq.y * rv.y;
  }
  block_finish<0>(ar, 0.0, partials, counter, aout, 0.0);
}

// Lanczos: r -= a q (a device scalar); b = ||r|| -> bout[0]
template <typename VT>
__global__ void __launch_bounds__(256) k_lz_axpy_nrm(const double* __restrict__ adev, const VT* __restrict__ q,
                                                     VT* __restrict__ r, int64_t n, double* partials,
                                                     unsigned int* counter, double* bout) {
  const double a = adev[0];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 rv = VTr<VT>::ld(r + i);
    const double2 qq = VTr<VT>::ld(q + i);
    rv.x -= a * qq.x;
    rv.y -= a * qq.y;
    VTr<VT>::st(r + i, rv);
    rv = VTr<VT>::rnd(rv);  // the stored value, without a reload
    s += rv.x * rv.x + rv.y * rv.y;
  }
  block_finish<0>(s, 0.0, partials, counter, bout, 1.0);
}

// q = r / b (b device scalar)
template <typename VT>
__global__ void __launch_bounds__(256) k_lz_div(const VT* __restrict__ r, const double* __restrict__ bdev,
                                                VT* __restrict__ q, int64_t n) {
  const double b = bdev[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = VTr<VT>::ld(r + i);
    VTr<VT>::st(q + i, make_double2(v.x / b, v.y / b));
  }
}

// ||x||_2 (FP64 accumulation)
template <typename VT>
__global__ void __launch_bounds__(256) k_nrm2(const VT* __restrict__ x, int64_t n, double* partials,
                                              unsigned int* counter, double* out) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = VTr<VT>::ld(x + i);
    s += v.x * v.x + v.y * v.y;
  }
  block_finish<0>(s, 0.0, partials, counter, out, 1.0);
}

// out[j] = sum_i c[i] Q[i, j]  (Q: m x n rows), c complex double
template <typename VT>
__global__ void __launch_bounds__(256) k_lincomb(const VT* __restrict__ Q, int64_t m, int64_t n,
                                                 const double2* __restrict__ c, VT* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t i = 0; i < m; ++i) {
      const double2 q = VTr<VT>::ld(Q + i * n + j);
      if constexpr (VTr<VT>::cplx) acc = cadd(acc, cmul(c[i], q));
      else acc.x += c[i].x * q.x;
    }
    VTr<VT>::st(out + j, acc);
  }
}

// scratch for reductions: partials (2*RED_BLOCKS doubles) + counter
struct RedScratch {
  double* partials = nullptr;
  unsigned int* counter = nullptr;
};

// partials (2 * RED_BLOCKS * 8 doubles) and 8 tickets; the tickets start at
// zero and every reduction kernel resets its own after the last block
static int alloc_red_scratch(RedScratch& R) {
  QSX_CHECK_HIP(hipMalloc(&R.partials, sizeof(double) * 2 * RED_BLOCKS * 8));
  QSX_CHECK_HIP(hipMalloc(&R.counter, sizeof(unsigned int) * 8));
  QSX_CHECK_HIP(hipMemset(R.counter, 0, sizeof(unsigned int) * 8));
  return QSX_OK;
}

// One scratch per (device, stream): reductions queued on different streams
// never share partials or tickets.  Launches on one stream run in order, so
// they may share it.  Kept until the process ends (64 KiB each).
static int get_red_scratch(RedScratch& R, hipStream_t st) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, RedScratch> per_stream;
  int cur = 0;
  QSX_CHECK_HIP(hipGetDevice(&cur));
  std::lock_guard<std::mutex> lk(mu);
  auto it = per_stream.find({cur, st});
  if (it == per_stream.end()) {
    RedScratch S;
    const int rc = alloc_red_scratch(S);
    if (rc) return rc;
    it = per_stream.emplace(std::make_pair(cur, st), S).first;
  }
  R = it->second;
  return QSX_OK;
}

// Blocks of a reduction: about eight elements per thread, at most
// RED_BLOCKS.  Every block pays a device-scope fence before its partial is
// counted, so small vectors use few blocks (measured on the L=28 Lanczos
// recurrence: 176 blocks instead of 512 for 361,270 elements, 11.3k -> 12.7k
// steps/s).
static int red_grid(int64_t n) {
  int64_t g = qsx_grid(n, 256 * 8);
  return (int)(g < RED_BLOCKS ? (g < 1 ? 1 : g) : RED_BLOCKS);
}

#define QSX_VT_SWITCH(dt, BODY)                                   \
  switch (dt) {                                                   \
    case QSX_F32: { using VT = float; BODY; } break;              \
    case QSX_F64: { using VT = double; BODY; } break;             \
    case QSX_C64: { using VT = float2; BODY; } break;             \
    case QSX_C128: { using VT = double2; BODY; } break;           \
    default: qsx_set_error("unknown dtype"); return QSX_ERR_ARG;  \
  }

// out[0] + i out[1] = sum conj(x) y   (device output, 2 doubles)
extern "C" int qsx_vdot(int dtype, const void* x, const void* y, int64_t n, double* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  RedScratch R;
  int rc = get_red_scratch(R, st);
  if (rc) return rc;
  QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_vdot<VT>, dim3(red_grid(n)), dim3(256), 0, st, (const VT*)x, (const VT*)y, n,
                                          R.partials, R.counter, out));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_nrm2(int dtype, const void* x, int64_t n, double* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  RedScratch R;
  int rc = get_red_scratch(R, st);
  if (rc) return rc;
  QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_nrm2<VT>, dim3(red_grid(n)), dim3(256), 0, st, (const VT*)x, n, R.partials,
                                          R.counter, out));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// y += a x ; a = (a_re, a_im) or, when sdev != NULL, sgn*sdev[0] (+i sgn*sdev[1] if cplx_scale)
extern "C" int qsx_axpy(int dtype, const void* x, void* y, int64_t n, double a_re, double a_im, const double* sdev,
                        double sgn, int cplx_scale, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_axpy_dev<VT>, dim3(red_grid(n) * 4), dim3(256), 0, st, (const VT*)x, (VT*)y, n,
                                          sdev, sgn, make_double2(a_re, a_im), cplx_scale));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// One Lanczos recurrence update after r = A q_i:
//   r -= b_prev v ; a_i = Re<q_i, r> ; r -= a_i q_i ; b_i = ||r||
// (src/quspin/tools/lanczos/_lanczos_utils.py:191-201).  b_prev / v may be
// NULL for the first step.  a_out, b_out: device doubles (2 slots each).
extern "C" int qsx_lanczos_update(int dtype, const void* v, const double* b_prev, const void* q, void* r, int64_t n,
                                  double* a_out, double* b_out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  RedScratch R;
  int rc = get_red_scratch(R, st);
  if (rc) return rc;
  QSX_VT_SWITCH(dtype, {
    hipLaunchKernelGGL(k_lz_axpy_dot<VT>, dim3(red_grid(n)), dim3(256), 0, st, (const VT*)v, b_prev, (const VT*)q,
                       (VT*)r, n, R.partials, R.counter, a_out);
    hipLaunchKernelGGL(k_lz_axpy_nrm<VT>, dim3(red_grid(n)), dim3(256), 0, st, (const double*)a_out, (const VT*)q,
                       (VT*)r, n, R.partials + 2 * RED_BLOCKS, R.counter + 1, b_out);
  });
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// Lanczos recurrence without fences: a kernel writes one partial per block
// (plain stores); the next kernel, after the launch boundary, sums them in a
// fixed order on every block (the same value everywhere, deterministic) and
// block 0 records the scalar.
// ---------------------------------------------------------------------------
// Lanczos recurrence grid: about two elements per lane, at most 1,024 blocks
// (the partial slots), so a 361 k vector spreads over all CUs.
static int lz_grid(int64_t n) {
  const int64_t g = (n + 511) / 512;
  return (int)(g < 1 ? 1 : (g > 1024 ? 1024 : g));
}

// b = sqrt(sum of the norm partials) -> bout; q = r / b
template <typename VT>
__global__ void __launch_bounds__(256) k_lzr_div(const VT* __restrict__ r, const double* __restrict__ pnrm, int nb,
                                                 VT* __restrict__ q, int64_t n, double* __restrict__ bout) {
  const double b = sqrt(sum_partials(pnrm, nb));
  if (blockIdx.x == 0 && threadIdx.x == 0) bout[0] = b;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = VTr<VT>::ld(r + i);
    VTr<VT>::st(q + i, make_double2(v.x / b, v.y / b));
  }
}

// r -= b_prev v (b_prev device scalar, or none); partial of Re <q, r>
template <typename VT>
__global__ void __launch_bounds__(256) k_lzr_dot(const VT* __restrict__ v, const double* __restrict__ bprev,
                                                 const VT* __restrict__ q, VT* __restrict__ r, int64_t n,
                                                 double* __restrict__ pdot) {
  const double b = bprev ? bprev[0] : 0.0;
  double ar = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 rv = VTr<VT>::ld(r + i);
    if (bprev) {
      const double2 vv = VTr<VT>::ld(v + i);
      rv.x -= b * vv.x;
      rv.y -= b * vv.y;
      VTr<VT>::st(r + i, rv);
      rv = VTr<VT>::rnd(rv);
    }
    const double2 qq = VTr<VT>::ld(q + i);
    ar += qq.x * rv.x + qq.y * rv.y;
  }
  block_partial(ar, pdot);
}

// a = sum of the dot partials -> aout; r -= a q; partial of ||r||^2
template <typename VT>
__global__ void __launch_bounds__(256) k_lzr_nrm(const double* __restrict__ pdot, int nb, const VT* __restrict__ q,
                                                 VT* __restrict__ r, int64_t n, double* __restrict__ aout,
                                                 double* __restrict__ pnrm) {
  const double a = sum_partials(pdot, nb);
  if (blockIdx.x == 0 && threadIdx.x == 0) aout[0] = a;
  double s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 rv = VTr<VT>::ld(r + i);
    const double2 qq = VTr<VT>::ld(q + i);
    rv.x -= a * qq.x;
    rv.y -= a * qq.y;
    VTr<VT>::st(r + i, rv);
    rv = VTr<VT>::rnd(rv);
    s2 += rv.x * rv.x + rv.y * rv.y;
  }
  block_partial(s2, pnrm);
}

__global__ void __launch_bounds__(256) k_lzr_fin(const double* __restrict__ pnrm, int nb, double* __restrict__ bout) {
  const double b = sqrt(sum_partials(pnrm, nb));
  if (threadIdx.x == 0) bout[0] = b;
}

// q = r / b (device scalar)
extern "C" int qsx_lanczos_div(int dtype, const void* r, const double* b, void* q, int64_t n, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_lz_div<VT>, dim3(red_grid(n) * 4), dim3(256), 0, st, (const VT*)r, b, (VT*)q, n));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// out = sum_i c_i Q[i]  (c: m complex doubles, device)
extern "C" int qsx_lincomb(int dtype, const void* Q, int64_t m, int64_t n, const double* c, void* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_lincomb<VT>, dim3(red_grid(n) * 4), dim3(256), 0, st, (const VT*)Q, m, n,
                                          (const double2*)c, (VT*)out));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// CSR trace and 1-norm
// ---------------------------------------------------------------------------
template <typename MT>
__global__ void __launch_bounds__(256) k_csr_trace(const void* indptr, const void* indices, const MT* __restrict__ data,
                                                   int idx64, int64_t Ns, double* partials, unsigned int* counter,
                                                   double* out) {
  double tr = 0.0, ti = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < Ns; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k0 = ld_idx(indptr, idx64, r), k1 = ld_idx(indptr, idx64, r + 1);
    for (int64_t k = k0; k < k1; ++k)
      if (ld_idx(indices, idx64, k) == r) { const double2 v = MTr<MT>::ld(data + k); tr += v.x; ti += v.y; }
  }
  block_finish<0>(tr, ti, partials, counter, out, 0.0);
}

template <typename MT>
__global__ void __launch_bounds__(256) k_csr_colabs(const void* indptr, const void* indices, const MT* __restrict__ data,
                                                    int idx64, int64_t Ns, double2 shift, const double* __restrict__ xdiag,
                                                    double2 xs, double* __restrict__ colsum) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < Ns; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k0 = ld_idx(indptr, idx64, r), k1 = ld_idx(indptr, idx64, r + 1);
    // diagonal addition of this row: shift + xs * xdiag[r]
    double2 dadd = shift;
    if (xdiag) dadd = cadd(dadd, make_double2(xs.x * xdiag[r], xs.y * xdiag[r]));
    bool has_diag = false;
    for (int64_t k = k0; k < k1; ++k) {
      const int64_t c = ld_idx(indices, idx64, k);
      double2 v = MTr<MT>::ld(data + k);
      if (c == r) { v = cadd(v, dadd); has_diag = true; }
      atomicAdd(colsum + c, sqrt(v.x * v.x + v.y * v.y));
    }
    if (!has_diag) atomicAdd(colsum + r, sqrt(dadd.x * dadd.x + dadd.y * dadd.y));
  }
}

__global__ void __launch_bounds__(256) k_max_reduce(const double* __restrict__ x, int64_t n, double* partials,
                                                    unsigned int* counter, double* out) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, x[i]);
  block_finish<1>(m, 0.0, partials, counter, out, 0.0);
}

// sum of n (re, im) pairs
__global__ void __launch_bounds__(256) k_sum_c(const double2* __restrict__ x, int64_t n, double* partials,
                                               unsigned int* counter, double* out) {
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    a += x[i].x;
    b += x[i].y;
  }
  block_finish<0>(a, b, partials, counter, out, 0.0);
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

#define QSX_MT_SWITCH(dt, BODY)                                   \
  switch (dt) {                                                   \
    case QSX_F32: { using MT = float; BODY; } break;              \
    case QSX_F64: { using MT = double; BODY; } break;             \
    case QSX_C64: { using MT = float2; BODY; } break;             \
    case QSX_C128: { using MT = double2; BODY; } break;           \
    default: qsx_set_error("unknown dtype"); return QSX_ERR_ARG;  \
  }

// op 0: out[0] + i out[1] = sum of the n (re, im) pairs x; op 1: out[0] =
// max of the n doubles x (device output, deterministic order).
extern "C" int qsx_reduce(int op, const double* x, int64_t n, double* out, void* stream) {
  if ((op != 0 && op != 1) || n < 0 || !out || (n > 0 && !x)) { qsx_set_error("qsx_reduce: bad arguments"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  RedScratch R;
  int rc = get_red_scratch(R, st);
  if (rc) return rc;
  if (op == 0)
    hipLaunchKernelGGL(k_sum_c, dim3(red_grid(n)), dim3(256), 0, st, (const double2*)x, n, R.partials, R.counter, out);
  else
    hipLaunchKernelGGL(k_max_reduce, dim3(red_grid(n)), dim3(256), 0, st, x, n, R.partials, R.counter, out);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// out (device, 2 doubles) = trace
extern "C" int qsx_csr_trace(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                             int64_t Ns, double* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  RedScratch R;
  int rc = get_red_scratch(R, st);
  if (rc) return rc;
  QSX_MT_SWITCH(mat_dtype, hipLaunchKernelGGL(k_csr_trace<MT>, dim3(red_grid(Ns)), dim3(256), 0, st, indptr, indices,
                                              (const MT*)data, idx64, Ns, R.partials, R.counter, out));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// out (device) = max_j sum_i |A_ij + shift delta_ij| ; colsum: Ns device doubles (scratch)
extern "C" int qsx_csr_1_norm_diag(const void* indptr, const void* indices, const void* data, int idx64,
                                   int mat_dtype, int64_t Ns, double shift_re, double shift_im, const double* xdiag,
                                   double xs_re, double xs_im, double* colsum, double* out, void* stream);

extern "C" int qsx_csr_1_norm(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                              int64_t Ns, double shift_re, double shift_im, double* colsum, double* out,
                              void* stream) {
  return qsx_csr_1_norm_diag(indptr, indices, data, idx64, mat_dtype, Ns, shift_re, shift_im, nullptr, 0.0, 0.0,
                             colsum, out, stream);
}

// ||A + shift I + xs diag(xdiag)||_1 (xdiag: Ns device doubles or null): the
// 1-norm of a time-dependent H(t) = A + f(t) D whose dynamic part is diagonal.
extern "C" int qsx_csr_1_norm_diag(const void* indptr, const void* indices, const void* data, int idx64,
                                   int mat_dtype, int64_t Ns, double shift_re, double shift_im, const double* xdiag,
                                   double xs_re, double xs_im, double* colsum, double* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  RedScratch R;
  int rc = get_red_scratch(R, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill_f64, dim3(red_grid(Ns)), dim3(256), 0, st, colsum, Ns, 0.0);
  QSX_MT_SWITCH(mat_dtype, hipLaunchKernelGGL(k_csr_colabs<MT>, dim3(red_grid(Ns) * 4), dim3(256), 0, st, indptr,
                                              indices, (const MT*)data, idx64, Ns, make_double2(shift_re, shift_im),
                                              xdiag, make_double2(xs_re, xs_im), colsum));
  hipLaunchKernelGGL(k_max_reduce, dim3(red_grid(Ns)), dim3(256), 0, st, colsum, Ns, R.partials, R.counter, out);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// diagonal of a CSR (duplicates summed; 0 where a row has no diagonal entry)
template <typename MT>
__global__ void __launch_bounds__(256) k_csr_diag(const void* indptr, const void* indices, const MT* __restrict__ data,
                                                  int idx64, int64_t Ns, double2* __restrict__ dg) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < Ns; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k0 = ld_idx(indptr, idx64, r), k1 = ld_idx(indptr, idx64, r + 1);
    double2 d = make_double2(0.0, 0.0);
    for (int64_t k = k0; k < k1; ++k)
      if (ld_idx(indices, idx64, k) == r) d = cadd(d, MTr<MT>::ld(data + k));
    dg[r] = d;
  }
}

// Column data of A for 1-norms of A + shifted diagonals: colabs[j] =
// sum_i |A_ij| and dg[j] = A_jj (Ns device doubles and Ns double2).
extern "C" int qsx_csr_col_split(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                                 int64_t Ns, double* colabs, double* dg, void* stream) {
  if (Ns < 0 || (Ns > 0 && (!indptr || !colabs || !dg))) { qsx_set_error("qsx_csr_col_split: bad arguments"); return QSX_ERR_ARG; }
  if (Ns == 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_f64, dim3(red_grid(Ns)), dim3(256), 0, st, colabs, Ns, 0.0);
  QSX_MT_SWITCH(mat_dtype, hipLaunchKernelGGL(k_csr_colabs<MT>, dim3(red_grid(Ns) * 4), dim3(256), 0, st, indptr,
                                              indices, (const MT*)data, idx64, Ns, make_double2(0.0, 0.0), nullptr,
                                              make_double2(0.0, 0.0), colabs));
  QSX_MT_SWITCH(mat_dtype, hipLaunchKernelGGL(k_csr_diag<MT>, dim3(red_grid(Ns)), dim3(256), 0, st, indptr, indices,
                                              (const MT*)data, idx64, Ns, (double2*)dg));
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// Exact Taylor partitions of a block of time steps, on the device.  For step
// record k, ||A + dscale_k D - mu_k I||_1 = max_j (colabs_j - |dg_j| +
// |dg_j + dscale_k D_j - mu_k|) from the column data of A, then (m*, s) by
// the theta branch of _fragment_3_1 (scipy _expm_multiply.py:503-558 as
// EMP.html:377-385 calls it, n0 = 1, ell = 2): s = ceil(|a| norm / theta_m),
// the smallest m * s over the theta table, first minimum kept.  A step whose
// |a| norm exceeds the branch bound is flagged in fail[k] (its partition
// needs the power-norm estimates of the other branch).
// ---------------------------------------------------------------------------
constexpr int PX_KB = 8;  // steps per thread of the norm pass

__global__ void k_part_clear(unsigned long long* __restrict__ nmax, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) nmax[k] = 0ull;
}

__global__ void __launch_bounds__(256) k_part_norm(const double* __restrict__ colabs, const double2* __restrict__ dg,
                                                   const double* __restrict__ D, int64_t N, const ExpmSec* __restrict__ es,
                                                   int nsteps, unsigned long long* __restrict__ nmax) {
  const int k0 = blockIdx.y * PX_KB;
  double2 dsc[PX_KB], mu[PX_KB];
  double best[PX_KB];
#pragma unroll
  for (int b = 0; b < PX_KB; ++b) {
    const int k = k0 + b < nsteps ? k0 + b : nsteps - 1;
    dsc[b] = es[k].dscale;
    mu[b] = es[k].mu;
    best[b] = 0.0;
  }
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (int64_t)gridDim.x * blockDim.x) {
    const double2 d = dg[j];
    const double off = fmax(colabs[j] - hypot(d.x, d.y), 0.0);
    const double dj = D ? D[j] : 0.0;
#pragma unroll
    for (int b = 0; b < PX_KB; ++b) {
      const double re = d.x + dsc[b].x * dj - mu[b].x, im = d.y + dsc[b].y * dj - mu[b].y;
      best[b] = fmax(best[b], off + hypot(re, im));
    }
  }
  __shared__ double red[PX_KB][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int b = 0; b < PX_KB; ++b) {
    double v = best[b];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if (lane == 0) red[b][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < PX_KB && k0 + (int)threadIdx.x < nsteps) {
    const int b = threadIdx.x;
    const double v = fmax(fmax(red[b][0], red[b][1]), fmax(red[b][2], red[b][3]));
    // non-negative doubles order as their bit patterns
    atomicMax(nmax + k0 + b, (unsigned long long)__double_as_longlong(v));
  }
}

__global__ void k_part_theta(ExpmSec* __restrict__ es, int nsteps, const unsigned long long* __restrict__ nmax,
                             const double* __restrict__ theta, const int* __restrict__ mth, int nth, double bound,
                             int* __restrict__ fail) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsteps) return;
  const double nrm = __longlong_as_double((long long)nmax[k]);
  const double one = hypot(es[k].a.x, es[k].a.y) * nrm;
  int bm = 0, bs = 1, f = 0;
  if (one > 0.0) {
    if (one <= bound) {
      long long best_m = -1, best_s = 0;
      for (int i = 0; i < nth; ++i) {
        const long long sv = (long long)ceil(one / theta[i]);
        if (best_m < 0 || (long long)mth[i] * sv < best_m * best_s) {
          best_m = mth[i];
          best_s = sv;
        }
      }
      bm = (int)best_m;
      bs = (int)best_s;
    } else {
      bm = 1;
      f = 1;
    }
  }
  es[k].m = bm;
  es[k].s = bs;
  fail[k] = f;
}

// es: nsteps ExpmSec records (a, mu, dscale set; m and s written); colabs/dg
// from qsx_csr_col_split (or the matrix-free column pass), D the dynamic
// diagonal (or null); theta/mth: the nth theta_m values and their m, in the
// order the selection visits them; bound: the branch bound on |a| norm;
// nmax: nsteps device words of scratch; fail: nsteps ints.
extern "C" int qsx_partition_exact(const double* colabs, const double* dg, const double* D, int64_t N, void* es,
                                   int nsteps, const double* theta, const int* mth, int nth, double bound,
                                   unsigned long long* nmax, int* fail, void* stream) {
  if (!es || nsteps < 1 || !theta || !mth || nth < 1 || !nmax || !fail || N < 0 || (N > 0 && (!colabs || !dg))) {
    qsx_set_error("qsx_partition_exact: bad arguments");
    return QSX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_part_clear, dim3((nsteps + 255) / 256), dim3(256), 0, st, nmax, nsteps);
  if (N > 0) {
    const dim3 grid((unsigned)std::min<int64_t>(red_grid(N), 512), (unsigned)((nsteps + PX_KB - 1) / PX_KB));
    hipLaunchKernelGGL(k_part_norm, grid, dim3(256), 0, st, colabs, (const double2*)dg, D, N, (const ExpmSec*)es,
                       nsteps, nmax);
  }
  hipLaunchKernelGGL(k_part_theta, dim3((nsteps + 255) / 256), dim3(256), 0, st, (ExpmSec*)es, nsteps, nmax, theta,
                     mth, nth, bound, fail);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// Taylor-series expm: F <- e^{a(A - mu I)} ... e^{a mu} F, in place
// ---------------------------------------------------------------------------
// ctl layout (device doubles): [0] done flag (as int in the first 4 bytes),
// [1] tol, then normB[s*(m+1)], normF[s*(m+1)].
template <typename VT>
__global__ void __launch_bounds__(256) k_expm_start(const VT* __restrict__ F, VT* __restrict__ B, int64_t n,
                                                    int64_t nvec, double* __restrict__ normB0, int nsp) {
  double mx = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double rs = 0.0;
    for (int64_t j = 0; j < nvec; ++j) {
      const double2 v = VTr<VT>::ld(F + r * nvec + j);
      VTr<VT>::st(B + r * nvec + j, v);
      rs += VTr<VT>::cplx ? sqrt(v.x * v.x + v.y * v.y) : fabs(v.x);
    }
    mx = fmax(mx, rs);
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) slot_put(normB0, nsp, sm[0]);
}

// F *= eta ; B = F ; ||F|| -> normB0 of the next step ; done = 0
template <typename VT>
__global__ void __launch_bounds__(256) k_expm_scale(VT* __restrict__ F, VT* __restrict__ B, int64_t n, int64_t nvec,
                                                    double2 eta, double* __restrict__ normB0, ExpmCtl* ctl,
                                                    int nsp) {
  double mx = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double rs = 0.0;
    for (int64_t j = 0; j < nvec; ++j) {
      double2 v = VTr<VT>::ld(F + r * nvec + j);
      if constexpr (VTr<VT>::cplx) v = cmul(eta, v); else v.x *= eta.x;
      VTr<VT>::st(F + r * nvec + j, v);
      VTr<VT>::st(B + r * nvec + j, v);
      const double2 w = VTr<VT>::rnd(v);
      rs += VTr<VT>::cplx ? sqrt(w.x * w.x + w.y * w.y) : fabs(w.x);
    }
    mx = fmax(mx, rs);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 0;
  if (!normB0) return;
  __shared__ double sm[256];
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) slot_put(normB0, nsp, sm[0]);
}

__global__ void k_expm_init_ctl(ExpmCtl* ctl, double tol, double* slots, int64_t nslots) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (int64_t)gridDim.x * blockDim.x)
    slots[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) { ctl->done = 0; ctl->tol = tol; }
}

// Matrix-free epilogue: test previous term, F += B, norms.
__global__ void k_expm_check(ExpmCtl* ctl, const double* c1, const double* c2, const double* nf) {
  if (ctl->done) return;
  if (*c1 + *c2 <= ctl->tol * *nf) ctl->done = 1;
}

template <typename VT>
__global__ void __launch_bounds__(256) k_expm_epi(const VT* __restrict__ B, VT* __restrict__ F, int64_t n, int64_t nvec,
                                                  const ExpmCtl* ctl, double* normB, double* normF) {
  if (ctl->done) return;
  double mb = 0.0, mf = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double rb = 0.0, rf = 0.0;
    for (int64_t j = 0; j < nvec; ++j) {
      const double2 b = VTr<VT>::ld(B + r * nvec + j);
      const double2 f = VTr<VT>::rnd(cadd(VTr<VT>::ld(F + r * nvec + j), b));
      VTr<VT>::st(F + r * nvec + j, f);
      rb += VTr<VT>::cplx ? sqrt(b.x * b.x + b.y * b.y) : fabs(b.x);
      rf += VTr<VT>::cplx ? sqrt(f.x * f.x + f.y * f.y) : fabs(f.x);
    }
    mb = fmax(mb, rb);
    mf = fmax(mf, rf);
  }
  __shared__ double sb[256], sf[256];
  sb[threadIdx.x] = mb;
  sf[threadIdx.x] = mf;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + o]); sf[threadIdx.x] = fmax(sf[threadIdx.x], sf[threadIdx.x + o]); }
    __syncthreads();
  }
  if (threadIdx.x == 0) { atomic_max_nonneg(normB, sb[0]); atomic_max_nonneg(normF, sf[0]); }
}

struct ExpmPlan {
  int s, m;
  double tol;
  double2 a, mu;
  int64_t Ns, nvec;
  int vdt;
};

// doubles per norm slot in the CSR Taylor loops (their SpMV launches have one
// block per 256/G rows; the matrix-free loop's epilogue grid is small: 1)
constexpr int CSR_NSP = 8;

static int expm_ctl_alloc(const ExpmPlan& P, ExpmCtl** ctl, double** slots, hipStream_t st, int nsp = 1) {
  const int64_t nsl = (int64_t)P.s * (P.m + 1) * 2 * nsp;
  // slots, then (CSR loops) 2 * EPI_MAX_BLOCKS per-block partials
  QSX_CHECK_HIP(hipMallocAsync((void**)ctl, sizeof(ExpmCtl) + sizeof(double) * (nsl + 2 + 2 * EPI_MAX_BLOCKS), st));
  *slots = (double*)((char*)*ctl + sizeof(ExpmCtl));
  hipLaunchKernelGGL(k_expm_init_ctl, dim3(qsx_grid(nsl, 256) > 0 ? (unsigned)qsx_grid(nsl, 256) : 1u), dim3(256), 0, st,
                     *ctl, P.tol, *slots, nsl);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// CSR Taylor loop on a caller-provided control block: ctl (ExpmCtl, then
// slots and per-block partials) initialised in-kernel, so the block may be
// reused by every call.
static int expm_csr_loop(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                         int64_t Ns, int64_t nnz, const ExpmPlan& P, void* v, void* work, int group, ExpmCtl* ctl,
                         hipStream_t st) {
  const int s = P.s, m_star = P.m;
  const double a_re = P.a.x, a_im = P.a.y, mu_re = P.mu.x, mu_im = P.mu.y;
  const int vec_dtype = P.vdt;
  const int64_t nvec = P.nvec;
  const int nsp = CSR_NSP;
  double* slots = (double*)((char*)ctl + sizeof(ExpmCtl));
  const int64_t nsl = (int64_t)s * (m_star + 1) * 2 * nsp;
  hipLaunchKernelGGL(k_expm_init_ctl, dim3(qsx_grid(nsl, 256) > 0 ? (unsigned)qsx_grid(nsl, 256) : 1u), dim3(256), 0, st,
                     ctl, P.tol, slots, nsl);
  QSX_LAUNCH_CHECK();
  const int64_t stride = (int64_t)(m_star + 1) * nsp;  // slot (i, j) at normB + i * stride + j * nsp
  double* normB = slots;
  double* normF = slots + (int64_t)s * stride;
  size_t esz = (vec_dtype == QSX_F32) ? 4 : (vec_dtype == QSX_F64 || vec_dtype == QSX_C64) ? 8 : 16;
  void* Bb[2] = {work, (char*)work + esz * Ns * nvec};
  const int G = group > 0 ? group : auto_group(Ns, nnz);
  const unsigned rg = (unsigned)red_grid(Ns) * 4;
  // eta = exp(a mu / s)
  const double er = a_re * mu_re - a_im * mu_im, ei = a_re * mu_im + a_im * mu_re;
  const double mag = std::exp(er / s);
  const double2 eta = make_double2(mag * std::cos(ei / s), mag * std::sin(ei / s));
  QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL(k_expm_start<VT>, dim3(rg), dim3(256), 0, st, (const VT*)v, (VT*)Bb[0], Ns,
                                              nvec, normB, nsp));
  QSX_LAUNCH_CHECK();
  for (int i = 0; i < s; ++i) {
    for (int j = 0; j < m_star; ++j) {
      CsrArgs A;
      std::memset(&A, 0, sizeof(A));
      A.indptr = indptr; A.indices = indices; A.data = data; A.idx64 = idx64; A.Ns = Ns; A.nvec = nvec;
      A.x = Bb[j & 1]; A.y = Bb[(j + 1) & 1];
      const double cf = 1.0 / ((double)s * (double)(j + 1));
      A.a = make_double2(a_re * cf, a_im * cf);
      A.mu = P.mu;
      A.beta = 0;
      A.F = v;
      A.ctl = ctl;
      A.nsp = nsp;
      A.part = slots + 2 * (int64_t)s * stride + 2;
      A.normB = normB + i * stride + (j + 1) * nsp;
      A.normF = normF + i * stride + (j + 1) * nsp;
      if (j >= 1) {
        A.chk_c1 = normB + i * stride + (j - 1) * nsp;
        A.chk_c2 = normB + i * stride + j * nsp;
        A.chk_F = normF + i * stride + j * nsp;
      }
      const int rc = launch_csr<true>(A, mat_dtype, vec_dtype, G, st);
      if (rc) return rc;
    }
    double* nb0 = (i + 1 < s) ? normB + (i + 1) * stride : nullptr;
    QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL(k_expm_scale<VT>, dim3(rg), dim3(256), 0, st, (VT*)v, (VT*)Bb[0], Ns,
                                                nvec, eta, nb0, ctl, nsp));
    QSX_LAUNCH_CHECK();
  }
  return QSX_OK;
}

// bytes of the control block of an (s, m_star) loop: ExpmCtl, the norm
// slots, 2 spare doubles and the per-block norm partials of the term kernel
static int64_t expm_csr_ctl_bytes(int s, int m_star) {
  return (int64_t)sizeof(ExpmCtl) +
         (int64_t)sizeof(double) * ((int64_t)s * (m_star + 1) * 2 * CSR_NSP + 2 + 2 * EPI_MAX_BLOCKS);
}

static int expm_csr_check(int64_t Ns, int s, int m_star, int vec_dtype, double a_im, double mu_im, const char* who) {
  if (s < 1 || m_star < 0) { qsx_set_error(std::string(who) + ": bad partition"); return QSX_ERR_ARG; }
  const bool real = (vec_dtype == QSX_F32 || vec_dtype == QSX_F64);
  if (real && (a_im != 0.0 || mu_im != 0.0)) { qsx_set_error("real vectors with complex a or mu"); return QSX_ERR_COMPLEX; }
  (void)Ns;
  return QSX_OK;
}

// CSR Taylor loop: v (Ns, nvec) is F and is overwritten; work has 2*Ns*nvec.
extern "C" int qsx_expm_multiply_csr(const void* indptr, const void* indices, const void* data, int idx64,
                                     int mat_dtype, int64_t Ns, int64_t nnz, int s, int m_star, double tol,
                                     double mu_re, double mu_im, double a_re, double a_im, void* v, void* work,
                                     int vec_dtype, int64_t nvec, int group, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (Ns <= 0) return QSX_OK;
  int rc = expm_csr_check(Ns, s, m_star, vec_dtype, a_im, mu_im, "qsx_expm_multiply_csr");
  if (rc) return rc;
  ExpmPlan P{s, m_star, tol, make_double2(a_re, a_im), make_double2(mu_re, mu_im), Ns, nvec, vec_dtype};
  ExpmCtl* ctl = nullptr;
  QSX_CHECK_HIP(hipMallocAsync((void**)&ctl, (size_t)expm_csr_ctl_bytes(s, m_star), st));
  rc = expm_csr_loop(indptr, indices, data, idx64, mat_dtype, Ns, nnz, P, v, work, group, ctl, st);
  (void)hipFreeAsync(ctl, st);
  return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free form of qsx_expm_multiply_csr.  Its control block lives in
// the file-scope device array `scratch` (one copy per GPU, kept until the
// process ends); the host addresses it only through the pointer that
// hipGetSymbolAddress gives for the current device, fetched once per device.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU race.  The loop's init kernel clears what must start at zero on every
// call (no memset node, so the sequence can be captured).
// ---------------------------------------------------------------------------
constexpr int64_t QSX_NOALLOC_SCRATCH_BYTES = 4ll << 20;  // 4 MiB
static __device__ __align__(256) unsigned char scratch[QSX_NOALLOC_SCRATCH_BYTES];

static int noalloc_scratch(void** out) {
  constexpr int MAX_DEV = 64;
  static std::mutex mu;
  static void* ptrs[MAX_DEV] = {};
  int dev = 0;
  QSX_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEV) { qsx_set_error("device index out of range"); return QSX_ERR_UNSUPPORTED; }
  std::lock_guard<std::mutex> lk(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      qsx_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return QSX_ERR_HIP;
    }
    ptrs[dev] = p;
  }
  *out = ptrs[dev];
  return QSX_OK;
}

extern "C" int64_t qsx_expm_multiply_csr_noalloc_max_bytes(void) { return QSX_NOALLOC_SCRATCH_BYTES; }

extern "C" int qsx_expm_multiply_csr_noalloc(const void* indptr, const void* indices, const void* data, int idx64,
                                             int mat_dtype, int64_t Ns, int64_t nnz, int s, int m_star, double tol,
                                             double mu_re, double mu_im, double a_re, double a_im, void* v,
                                             void* work, int vec_dtype, int64_t nvec, int group, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (Ns <= 0) return QSX_OK;
  int rc = expm_csr_check(Ns, s, m_star, vec_dtype, a_im, mu_im, "qsx_expm_multiply_csr_noalloc");
  if (rc) return rc;
  const int64_t need = expm_csr_ctl_bytes(s, m_star);
  if (need > QSX_NOALLOC_SCRATCH_BYTES) {
    qsx_set_error("qsx_expm_multiply_csr_noalloc: partition (s = " + std::to_string(s) + ", m = " +
                  std::to_string(m_star) + ") needs " + std::to_string(need) + " bytes of control block, above the " +
                  std::to_string(QSX_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
    return QSX_ERR_ARG;
  }
  void* p = nullptr;
  rc = noalloc_scratch(&p);
  if (rc) return rc;
  ExpmPlan P{s, m_star, tol, make_double2(a_re, a_im), make_double2(mu_re, mu_im), Ns, nvec, vec_dtype};
  return expm_csr_loop(indptr, indices, data, idx64, mat_dtype, Ns, nnz, P, v, work, group, (ExpmCtl*)p, st);
}

// Matrix-free Taylor loop: same recurrence with H.v from the fused kernel.
int qsx_hv_expm_term(void* hv_ctx, const void* x, void* y, double2 coef, double2 shift, const int* skip,
                     hipStream_t st);
int qsx_hv_expm_term_epi(void* hv_ctx, const void* x, void* y, double2 coef, double2 shift, const int* skip,
                         void* F, double* nB, double* nF, hipStream_t st);

extern "C" int qsx_expm_multiply_mf(void* hv_ctx, int64_t Ns, int s, int m_star, double tol, double mu_re,
                                    double mu_im, double a_re, double a_im, void* v, void* work, int vec_dtype,
                                    int64_t nvec, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (Ns <= 0) return QSX_OK;
  if (s < 1 || m_star < 0) { qsx_set_error("qsx_expm_multiply_mf: bad partition"); return QSX_ERR_ARG; }
  ExpmPlan P{s, m_star, tol, make_double2(a_re, a_im), make_double2(mu_re, mu_im), Ns, nvec, vec_dtype};
  ExpmCtl* ctl = nullptr;
  double* slots = nullptr;
  int rc = expm_ctl_alloc(P, &ctl, &slots, st);
  if (rc) return rc;
  const int64_t stride = (int64_t)(m_star + 1);
  double* normB = slots;
  double* normF = slots + (int64_t)s * stride;
  size_t esz = (vec_dtype == QSX_F32) ? 4 : (vec_dtype == QSX_F64 || vec_dtype == QSX_C64) ? 8 : 16;
  void* Bb[2] = {work, (char*)work + esz * Ns * nvec};
  const unsigned rg = (unsigned)red_grid(Ns) * 4;
  const double er = a_re * mu_re - a_im * mu_im, ei = a_re * mu_im + a_im * mu_re;
  const double mag = std::exp(er / s);
  const double2 eta = make_double2(mag * std::cos(ei / s), mag * std::sin(ei / s));
  bool fuse = nvec == 1;
  QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL(k_expm_start<VT>, dim3(rg), dim3(256), 0, st, (const VT*)v, (VT*)Bb[0], Ns,
                                              nvec, normB, 1));
  QSX_LAUNCH_CHECK();
  for (int i = 0; i < s; ++i) {
    for (int j = 0; j < m_star; ++j) {
      if (j >= 1) {
        hipLaunchKernelGGL(k_expm_check, dim3(1), dim3(1), 0, st, ctl, normB + i * stride + j - 1,
                           normB + i * stride + j, normF + i * stride + j);
      }
      const double cf = 1.0 / ((double)s * (double)(j + 1));
      const double2 coef = make_double2(a_re * cf, a_im * cf);
      // shift = -coef * mu
      const double2 shift = make_double2(-(coef.x * mu_re - coef.y * mu_im), -(coef.x * mu_im + coef.y * mu_re));
      if (fuse) {
        // the two-pass split product writes the term and updates F and the
        // norms in its strip pass (no separate epilogue pass)
        rc = qsx_hv_expm_term_epi(hv_ctx, Bb[j & 1], Bb[(j + 1) & 1], coef, shift, &ctl->done, v,
                                  normB + i * stride + j + 1, normF + i * stride + j + 1, st);
        if (rc == QSX_OK) continue;
        if (rc != QSX_ERR_UNSUPPORTED) return rc;
        fuse = false;
      }
      rc = qsx_hv_expm_term(hv_ctx, Bb[j & 1], Bb[(j + 1) & 1], coef, shift, &ctl->done, st);
      if (rc) return rc;
      QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL(k_expm_epi<VT>, dim3(rg), dim3(256), 0, st, (const VT*)Bb[(j + 1) & 1],
                                                  (VT*)v, Ns, nvec, ctl, normB + i * stride + j + 1,
                                                  normF + i * stride + j + 1));
      QSX_LAUNCH_CHECK();
    }
    double* nb0 = (i + 1 < s) ? normB + (i + 1) * stride : nullptr;
    QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL(k_expm_scale<VT>, dim3(rg), dim3(256), 0, st, (VT*)v, (VT*)Bb[0], Ns,
                                                nvec, eta, nb0, ctl, 1));
    QSX_LAUNCH_CHECK();
  }
  (void)hipFreeAsync(ctl, st);
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// plain device copy (16-byte vectors), the reference point for HBM counters
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" int qsx_copy(const void* src, void* dst, int64_t bytes, void* stream) {
  if (bytes % 16) { qsx_set_error("qsx_copy: size must be a multiple of 16 bytes"); return QSX_ERR_ARG; }
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3(2048), dim3(256), 0, (hipStream_t)stream, (const uint4*)src, (uint4*)dst, n);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// Time steps with device records: for t = 0..nsteps-1, v <- e^{a_t (A +
// dscale_t D)} v, partition / a / mu / tol / dscale read from es[t]
// ---------------------------------------------------------------------------
// F *= eta(record) ; B = F ; ||F|| -> normB0 of the next stage ; done = 0 (stage i of the record)
template <typename FT, typename VT>
__global__ void __launch_bounds__(256) k_expm_scale_rec(FT* __restrict__ F, VT* __restrict__ B, int64_t n,
                                                        int64_t nvec, const ExpmSec* rec, int stage,
                                                        double* __restrict__ normB0, ExpmCtl* ctl, int nsp) {
  const ExpmSec E = *rec;
  if (stage >= E.s) return;
  const double2 eta = expm_eta(E);
  double mx = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double rs = 0.0;
    for (int64_t j = 0; j < nvec; ++j) {
      double2 v = VTr<FT>::ld(F + r * nvec + j);
      if constexpr (VTr<FT>::cplx) v = cmul(eta, v); else v.x *= eta.x;
      VTr<FT>::st(F + r * nvec + j, v);
      VTr<VT>::st(B + r * nvec + j, v);
      const double2 w = VTr<FT>::rnd(v);
      rs += VTr<FT>::cplx ? sqrt(w.x * w.x + w.y * w.y) : fabs(w.x);
    }
    mx = fmax(mx, rs);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 0;
  if (stage + 1 >= E.s) return;
  __shared__ double sm[256];
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) slot_put(normB0, nsp, sm[0]);
}

// B = F (rounded to the term dtype), ||F|| -> normB0
template <typename FT, typename VT>
__global__ void __launch_bounds__(256) k_expm_start_mixed(const FT* __restrict__ F, VT* __restrict__ B, int64_t n,
                                                          int64_t nvec, double* __restrict__ normB0, int nsp) {
  double mx = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double rs = 0.0;
    for (int64_t j = 0; j < nvec; ++j) {
      const double2 v = VTr<FT>::ld(F + r * nvec + j);
      VTr<VT>::st(B + r * nvec + j, v);
      rs += VTr<FT>::cplx ? sqrt(v.x * v.x + v.y * v.y) : fabs(v.x);
    }
    mx = fmax(mx, rs);
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) slot_put(normB0, nsp, sm[0]);
}

// A stage that ran all m terms of its record without passing the truncation
// test flags the step.
__global__ void k_expm_tail(const ExpmSec* rec, int stage, const ExpmCtl* ctl, const double* nb, const double* nf,
                            int nsp, int* fail) {
  const ExpmSec E = *rec;
  if (stage >= E.s || E.m < 1 || ctl->done) return;
  const double c1 = slot_max(nb + (int64_t)(E.m - 1) * nsp, nsp), c2 = slot_max(nb + (int64_t)E.m * nsp, nsp);
  if (!(c1 + c2 <= E.tol * slot_max(nf + (int64_t)E.m * nsp, nsp))) *fail = 1;
}

// zero the norm slots and the done flag (graph-safe: no memset nodes)
__global__ void k_expm_reset(ExpmCtl* ctl, double* slots, int64_t nslots) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (int64_t)gridDim.x * blockDim.x)
    slots[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 0;
}

extern "C" int64_t qsx_evolve_csr_ctl_bytes(int s_max, int m_max) {
  return (int64_t)sizeof(ExpmCtl) +
         (int64_t)sizeof(double) * (2 * (int64_t)s_max * (m_max + 1) * CSR_NSP + 2 * EPI_MAX_BLOCKS);
}

namespace {
// A captured sequence, the object whose fields it holds (0: none; the key
// then names every array and size the sequence was recorded with) and the
// reduction scratch it owns (Lanczos captures).
struct GraphEntry {
  hipGraphExec_t ex = nullptr;
  uint64_t owner = 0;
  uint64_t seq = 0;
  RedScratch scratch;
};
using StepKey = std::tuple<uint64_t, const void*, const void*, const void*, int, int64_t, const void*, const void*,
                           const void*, const void*, const void*, const void*, int64_t, int64_t, int, int, int, int, int>;
std::map<StepKey, GraphEntry> g_step_graphs;
using LzKey = std::tuple<uint64_t, const void*, const void*, const void*, int, int, int64_t, int, const void*,
                         const void*, const void*, int64_t, int, int, int, int>;
std::map<LzKey, GraphEntry> g_lz_graphs;
hipStream_t g_cap_stream = nullptr;
std::mutex g_graph_mu;
std::atomic<uint64_t> g_next_owner{1};
uint64_t g_graph_seq = 0;
// captures kept at most; the oldest is dropped first (callers that pass new
// buffers on every call would otherwise grow the cache without bound)
constexpr size_t GRAPH_CACHE_MAX = 64;

void destroy_entry(GraphEntry& e) {
  if (e.ex) (void)hipGraphExecDestroy(e.ex);
  if (e.scratch.partials) (void)hipFree(e.scratch.partials);
  if (e.scratch.counter) (void)hipFree(e.scratch.counter);
  e = GraphEntry{};
}

template <class Map, class Pred>
size_t erase_if_entries(Map& m, Pred pred) {
  size_t n = 0;
  for (auto it = m.begin(); it != m.end();) {
    if (pred(it->second)) {
      destroy_entry(it->second);
      it = m.erase(it);
      ++n;
    } else {
      ++it;
    }
  }
  return n;
}

bool any_entry(uint64_t owner) {
  for (auto& kv : g_step_graphs)
    if (kv.second.owner == owner) return true;
  for (auto& kv : g_lz_graphs)
    if (kv.second.owner == owner) return true;
  return false;
}

// make room for one more capture (g_graph_mu held)
void evict_oldest() {
  while (g_step_graphs.size() + g_lz_graphs.size() >= GRAPH_CACHE_MAX) {
    uint64_t oldest = UINT64_MAX;
    for (auto& kv : g_step_graphs) oldest = kv.second.seq < oldest ? kv.second.seq : oldest;
    for (auto& kv : g_lz_graphs) oldest = kv.second.seq < oldest ? kv.second.seq : oldest;
    (void)hipDeviceSynchronize();  // a replay of it may still be queued
    auto old = [&](const GraphEntry& e) { return e.seq == oldest; };
    erase_if_entries(g_step_graphs, old);
    erase_if_entries(g_lz_graphs, old);
  }
}
}  // namespace

uint64_t qsx_owner_new() { return g_next_owner.fetch_add(1); }

void qsx_graph_drop_owner(uint64_t owner) {
  if (!owner) return;
  std::lock_guard<std::mutex> lk(g_graph_mu);
  if (!any_entry(owner)) return;
  (void)hipDeviceSynchronize();  // queued replays finish before their graphs go
  auto mine = [&](const GraphEntry& e) { return e.owner == owner; };
  erase_if_entries(g_step_graphs, mine);
  erase_if_entries(g_lz_graphs, mine);
}

// Drop every captured sequence.
extern "C" void qsx_graph_cache_clear() {
  std::lock_guard<std::mutex> lk(g_graph_mu);
  if (g_step_graphs.empty() && g_lz_graphs.empty()) return;
  (void)hipDeviceSynchronize();
  auto all = [](const GraphEntry&) { return true; };
  erase_if_entries(g_step_graphs, all);
  erase_if_entries(g_lz_graphs, all);
}

// Capture record(stream) once into a new cache entry of `cache` under `key`
// (g_graph_mu held by the caller).
template <class Map, class Key, class Rec>
static int capture_into(Map& cache, const Key& key, uint64_t owner, const RedScratch& scratch, Rec&& record,
                        typename Map::iterator& out) {
  evict_oldest();
  if (!g_cap_stream) QSX_CHECK_HIP(hipStreamCreateWithFlags(&g_cap_stream, hipStreamNonBlocking));
  QSX_CHECK_HIP(hipStreamBeginCapture(g_cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = record(g_cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(g_cap_stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) { qsx_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return QSX_ERR_HIP; }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { qsx_set_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return QSX_ERR_HIP; }
  GraphEntry ent;
  ent.ex = ex;
  ent.owner = owner;
  ent.seq = ++g_graph_seq;
  ent.scratch = scratch;
  out = cache.emplace(key, ent).first;
  return QSX_OK;
}

using TermFn = int (*)(const CsrArgs&, const void*, int, int, hipStream_t);

// The time-step loop for any expm term kernel: for t = 0..nsteps-1, stage i,
// term j: term(A) with A carrying vectors, record, slots and partials; the
// term kernel writes its per-block norm partials (part_blocks blocks at most).
int qsx_evolve_steps_with(int64_t Ns, int64_t nvec, const double* xdiag, int nsteps, const void* es, int s_max,
                          int m_max, void* v, void* work, int vec_dtype, int state_dtype, int* fail, void* ctl_work,
                          int64_t ctl_bytes, int64_t part_blocks, TermFn term, const void* term_ctx,
                          const StepOpKey& opk, int use_graph, hipStream_t st) {
  if (Ns < 0 || nsteps < 1 || !es || s_max < 1 || m_max < 1 || !v || !work || !fail || !ctl_work || nvec < 1 ||
      ctl_bytes < qsx_evolve_csr_ctl_bytes(s_max, m_max) || part_blocks > EPI_MAX_BLOCKS) {
    qsx_set_error("qsx_evolve_steps: bad arguments");
    return QSX_ERR_ARG;
  }
  if (Ns == 0) return QSX_OK;
  if (state_dtype < 0) state_dtype = vec_dtype;
  const bool mixed = state_dtype != vec_dtype;
  if (mixed && !((vec_dtype == QSX_C64 && state_dtype == QSX_C128) || (vec_dtype == QSX_F32 && state_dtype == QSX_F64))) {
    qsx_set_error("qsx_evolve_steps: the state dtype must be the term dtype or its FP64 counterpart");
    return QSX_ERR_ARG;
  }
  ExpmCtl* ctl = (ExpmCtl*)ctl_work;
  double* slots = (double*)((char*)ctl_work + sizeof(ExpmCtl));
  const int nsp = CSR_NSP;
  const int64_t stride = (int64_t)(m_max + 1) * nsp;  // slot (i, j) at normB + i * stride + j * nsp
  const int64_t nsl = 2 * (int64_t)s_max * stride;
  double* normB = slots;
  double* normF = slots + (int64_t)s_max * stride;
  size_t esz = (vec_dtype == QSX_F32) ? 4 : (vec_dtype == QSX_F64 || vec_dtype == QSX_C64) ? 8 : 16;
  void* Bb[2] = {work, (char*)work + esz * Ns * nvec};
  const unsigned rg = (unsigned)red_grid(Ns) * 4;
  const ExpmSec* E = (const ExpmSec*)es;

  auto record = [&](hipStream_t q) -> int {
    for (int t = 0; t < nsteps; ++t) {
      const ExpmSec* R = E + t;
      hipLaunchKernelGGL(k_expm_reset, dim3(16), dim3(256), 0, q, ctl, slots, nsl);
      QSX_LAUNCH_CHECK();
      if (!mixed) {
        QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL(k_expm_start<VT>, dim3(rg), dim3(256), 0, q, (const VT*)v,
                                                    (VT*)Bb[0], Ns, nvec, normB, nsp));
      } else if (vec_dtype == QSX_C64) {
        hipLaunchKernelGGL((k_expm_start_mixed<double2, float2>), dim3(rg), dim3(256), 0, q, (const double2*)v,
                           (float2*)Bb[0], Ns, nvec, normB, nsp);
      } else {
        hipLaunchKernelGGL((k_expm_start_mixed<double, float>), dim3(rg), dim3(256), 0, q, (const double*)v,
                           (float*)Bb[0], Ns, nvec, normB, nsp);
      }
      QSX_LAUNCH_CHECK();
      for (int i = 0; i < s_max; ++i) {
        for (int j = 0; j < m_max; ++j) {
          CsrArgs A;
          std::memset(&A, 0, sizeof(A));
          A.Ns = Ns;
          A.nvec = nvec;
          A.x = Bb[j & 1]; A.y = Bb[(j + 1) & 1];
          A.F = v;
          A.ctl = ctl;
          A.rec = R;
          A.stage = i;
          A.term = j;
          A.xdiag = xdiag;
          A.nsp = nsp;
          A.part = slots + nsl;
          A.normB = normB + i * stride + (j + 1) * nsp;
          A.normF = normF + i * stride + (j + 1) * nsp;
          if (j >= 1) {
            A.chk_c1 = normB + i * stride + (j - 1) * nsp;
            A.chk_c2 = normB + i * stride + j * nsp;
            A.chk_F = normF + i * stride + j * nsp;
          }
          int rc = term(A, term_ctx, vec_dtype, state_dtype, q);
          if (rc) return rc;
        }
        hipLaunchKernelGGL(k_expm_tail, dim3(1), dim3(1), 0, q, R, i, ctl, normB + i * stride, normF + i * stride,
                           nsp, fail + t);
        QSX_LAUNCH_CHECK();
        double* nb0 = (i + 1 < s_max) ? normB + (i + 1) * stride : normB;
        if (!mixed) {
          QSX_VT_SWITCH(vec_dtype, hipLaunchKernelGGL((k_expm_scale_rec<VT, VT>), dim3(rg), dim3(256), 0, q, (VT*)v,
                                                      (VT*)Bb[0], Ns, nvec, R, i, nb0, ctl, nsp));
        } else if (vec_dtype == QSX_C64) {
          hipLaunchKernelGGL((k_expm_scale_rec<double2, float2>), dim3(rg), dim3(256), 0, q, (double2*)v,
                             (float2*)Bb[0], Ns, nvec, R, i, nb0, ctl, nsp);
        } else {
          hipLaunchKernelGGL((k_expm_scale_rec<double, float>), dim3(rg), dim3(256), 0, q, (double*)v, (float*)Bb[0],
                             Ns, nvec, R, i, nb0, ctl, nsp);
        }
        QSX_LAUNCH_CHECK();
      }
    }
    return QSX_OK;
  };
  if (!use_graph) return record(st);
  // the key names every operator array, size and buffer the sequence is
  // recorded with; an operator object's fields are covered by its owner id
  const StepKey key{opk.owner, opk.p[0], opk.p[1], opk.p[2], opk.idx64, opk.nnz, xdiag, es, v, work, fail,
                    ctl_work, Ns, nvec, nsteps, s_max, m_max, opk.kind, vec_dtype * 8 + state_dtype};
  std::lock_guard<std::mutex> lk(g_graph_mu);
  auto it = g_step_graphs.find(key);
  if (it == g_step_graphs.end()) {
    const int rc = capture_into(g_step_graphs, key, opk.owner, RedScratch{}, record, it);
    if (rc) return rc;
  }
  QSX_CHECK_HIP(hipGraphLaunch(it->second.ex, st));
  return QSX_OK;
}

namespace {
struct CsrTermCtx {
  const void* indptr;
  const void* indices;
  const void* data;
  int idx64, mat_dtype, G;
  int64_t nnz;
};

int csr_term(const CsrArgs& A0, const void* ctx, int vdt, int fdt, hipStream_t q) {
  const CsrTermCtx* c = (const CsrTermCtx*)ctx;
  CsrArgs A = A0;
  A.indptr = c->indptr;
  A.indices = c->indices;
  A.data = c->data;
  A.idx64 = c->idx64;
  return launch_csr<true>(A, c->mat_dtype, vdt, c->G, q, fdt);
}
}  // namespace

extern "C" int qsx_evolve_csr_steps(const void* indptr, const void* indices, const void* data, int idx64,
                                    int mat_dtype, int64_t Ns, int64_t nnz, const double* xdiag, int nsteps,
                                    const void* es, int s_max, int m_max, void* v, void* work, int vec_dtype,
                                    int state_dtype, int64_t nvec, int* fail, void* ctl_work, int64_t ctl_bytes,
                                    int group, int use_graph, void* stream) {
  if (Ns > 0 && (!indptr || !indices || !data)) { qsx_set_error("qsx_evolve_csr_steps: bad arguments"); return QSX_ERR_ARG; }
  CsrTermCtx c{indptr, indices, data, idx64, mat_dtype, group > 0 ? group : auto_group(Ns, nnz), nnz};
  const StepOpKey key{0, {indptr, indices, data}, idx64, nnz, c.G * 16 + mat_dtype};
  return qsx_evolve_steps_with(Ns, nvec, xdiag, nsteps, es, s_max, m_max, v, work, vec_dtype, state_dtype, fail,
                               ctl_work, ctl_bytes, csr_blocks(Ns, c.G, true), csr_term, &c, key, use_graph,
                               (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Whole Lanczos recurrence on the device (_lanczos_utils.py:158-227): m steps
// of div / H.v / fused updates with alpha, beta kept in device memory, as one
// launch sequence that can be captured once and replayed.
// ---------------------------------------------------------------------------
extern "C" int qsx_hv_ctx_apply(void* X, const void* x, void* y, double a_off_re, double a_off_im,
                                double a_diag_re, double a_diag_im, int beta, void* stream);

extern "C" int qsx_lanczos_run(int op_kind, void* hv_ctx, const void* indptr, const void* indices, const void* data,
                               int idx64, int mat_dtype, int64_t nnz, int group, int dtype, int64_t n, int m, void* Q,
                               int store_Q, void* r, double* ab, int use_graph, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n < 1 || m < 1 || !Q || !r || !ab || (op_kind == 0 && !hv_ctx) || (op_kind == 1 && (!indptr || !data))) {
    qsx_set_error("qsx_lanczos_run: bad arguments");
    return QSX_ERR_ARG;
  }
  // reduction partials: the stream's scratch for a direct launch sequence,
  // a scratch of the capture's own for a captured one (allocated before the
  // capture begins and freed with it)
  RedScratch R;
  const size_t esz = (dtype == QSX_F32) ? 4 : (dtype == QSX_F64 || dtype == QSX_C64) ? 8 : 16;
  auto qi = [&](int i) -> void* {
    return (char*)Q + esz * (size_t)n * (size_t)(store_Q ? i : (i % 3));
  };
  auto apply = [&](const void* x, hipStream_t q) -> int {
    if (op_kind == 0) return qsx_hv_ctx_apply(hv_ctx, x, r, 1.0, 0.0, 1.0, 0.0, 0, q);
    return qsx_csr_matvec(indptr, indices, data, idx64, mat_dtype, n, nnz, x, r, dtype, 1, 1.0, 0.0, 0, group, q);
  };
  // per step: q_i = r / b_{i-1} (the norm partials of the previous step
  // reduced on every block), r = A q_i, r -= b_{i-1} q_{i-1} with the
  // partials of <q_i, r>, r -= a_i q_i with the partials of ||r||^2
  const int nb = lz_grid(n);
  auto record = [&](hipStream_t q) -> int {
    double* pdot = R.partials + 4 * RED_BLOCKS;  // 1,024 slots each
    double* pnrm = R.partials + 6 * RED_BLOCKS;
    for (int i = 0; i < m; ++i) {
      double* bp = ab + 4 * (int64_t)(i - 1) + 2;  // b_{i-1}
      if (i > 0) {
        QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_lzr_div<VT>, dim3(nb), dim3(256), 0, q, (const VT*)r, pnrm, nb,
                                                (VT*)qi(i), n, bp));
        QSX_LAUNCH_CHECK();
      }
      int e = apply(qi(i), q);
      if (e) return e;
      QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_lzr_dot<VT>, dim3(nb), dim3(256), 0, q,
                                              i > 0 ? (const VT*)qi(i - 1) : (const VT*)nullptr,
                                              i > 0 ? bp : (const double*)nullptr, (const VT*)qi(i), (VT*)r, n,
                                              pdot));
      QSX_LAUNCH_CHECK();
      QSX_VT_SWITCH(dtype, hipLaunchKernelGGL(k_lzr_nrm<VT>, dim3(nb), dim3(256), 0, q, (const double*)pdot, nb,
                                              (const VT*)qi(i), (VT*)r, n, ab + 4 * (int64_t)i, pnrm));
      QSX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_lzr_fin, dim3(1), dim3(256), 0, q, (const double*)pnrm, nb, ab + 4 * (int64_t)(m - 1) + 2);
    QSX_LAUNCH_CHECK();
    return QSX_OK;
  };
  if (!use_graph) {
    const int rc = get_red_scratch(R, st);
    if (rc) return rc;
    return record(st);
  }
  const uint64_t owner = op_kind == 0 ? qsx_hv_ctx_owner((const qsx_hv_ctx*)hv_ctx) : 0;
  const LzKey key{owner, indptr, indices, data, idx64, mat_dtype, nnz, group, Q, r, ab, n, m, store_Q, dtype,
                  op_kind};
  std::lock_guard<std::mutex> lk(g_graph_mu);
  auto it = g_lz_graphs.find(key);
  if (it == g_lz_graphs.end()) {
    int rc = alloc_red_scratch(R);
    if (rc == QSX_OK) rc = capture_into(g_lz_graphs, key, owner, R, record, it);
    if (rc) {
      if (R.partials) (void)hipFree(R.partials);
      if (R.counter) (void)hipFree(R.counter);
      return rc;
    }
  }
  QSX_CHECK_HIP(hipGraphLaunch(it->second.ex, st));
  return QSX_OK;
}
