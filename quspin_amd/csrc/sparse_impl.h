// CSR kernels shared by hamiltonian.dot and the Taylor-series expm.
#pragma once
#include "hv_impl.h"

template <typename MT> struct MTr;
template <> struct MTr<float> { static constexpr bool cplx = false; __device__ static double2 ld(const float* p) { return make_double2((double)*p, 0.0); } };
template <> struct MTr<double> { static constexpr bool cplx = false; __device__ static double2 ld(const double* p) { return make_double2(*p, 0.0); } };
template <> struct MTr<float2> { static constexpr bool cplx = true; __device__ static double2 ld(const float2* p) { float2 v = *p; return make_double2((double)v.x, (double)v.y); } };
template <> struct MTr<double2> { static constexpr bool cplx = true; __device__ static double2 ld(const double2* p) { return *p; } };


// Sum over the 256 lanes of a block, returned on every lane: a shuffle tree
// inside each wavefront, then the four wavefront sums in a fixed order (the
// same value on every block for the same inputs).
__device__ __forceinline__ double block_sum256(double v) {
  __shared__ double sh[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// Sum of nb per-block partials, computed identically by every block.
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int nb) {
  double a = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) a += part[k];
  return block_sum256(a);
}

// This block's partial of a grid-wide sum into part[blockIdx.x].
__device__ __forceinline__ void block_partial(double v, double* __restrict__ part) {
  const double r = block_sum256(v);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Atomic max of a non-negative double through its bit pattern.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ double cabs2(double2 z) { return sqrt(z.x * z.x + z.y * z.y); }

// A norm slot is a group of nsp doubles: writers take the atomic max into
// entry blockIdx % nsp (spreading the same-address atomics of large grids),
// readers take the max over the group.
__device__ __forceinline__ double slot_max(const double* p, int nsp) {
  double m = p[0];
  for (int i = 1; i < nsp; ++i) m = fmax(m, p[i]);
  return m;
}
__device__ __forceinline__ void slot_put(double* p, int nsp, double v) { atomic_max_nonneg(p + (blockIdx.x % nsp), v); }

// Taylor-loop control block (one per expm call).
struct ExpmCtl {
  int done;      // set when the truncation test of the current outer step passed
  int pad;
  double tol;
  // normB[i*(m+1)+j], normF[i*(m+1)+j] follow in device memory
};

// Per-step (and per-sector) Taylor-loop record, filled on the host or by the
// partition predictor on the device.
struct ExpmSec {
  double2 a;       // e^{aA}
  double2 mu;      // tr(A)/N
  double2 dscale;  // factor of the extra diagonal (time-dependent term), if any
  double tol;
  int s, m;        // partition (s stages of at most m terms)
  int pad[2];
};

// e^{a mu / s}
__device__ __forceinline__ double2 expm_eta(const ExpmSec& E) {
  const double er = (E.a.x * E.mu.x - E.a.y * E.mu.y) / E.s, ei = (E.a.x * E.mu.y + E.a.y * E.mu.x) / E.s;
  const double mag = exp(er);
  return make_double2(mag * cos(ei), mag * sin(ei));
}

// The operator a captured step sequence was recorded from: the owner id of
// an object whose fields the term kernel holds (0 for none), its arrays and
// the sizes and kind that select the launches.
struct StepOpKey {
  uint64_t owner;
  const void* p[3];
  int idx64;
  int64_t nnz;
  int kind;
};

struct CsrArgs {
  const void* indptr;   // int32 or int64
  const void* indices;  // int32 or int64
  const void* data;
  int idx64;
  int64_t Ns;
  int64_t nvec;
  const void* x;
  void* y;
  double2 a;            // y = a*(A x - mu x) [+ y if beta]
  double2 mu;
  int beta;
  // fused Taylor epilogue (expm): F += y, norms, early exit
  void* F;
  ExpmCtl* ctl;
  double* normB;        // slot for ||y||_inf (this term)
  double* normF;        // slot for ||F||_inf (this term)
  const double* chk_c1; // previous test inputs: c1, c2, ||F|| (null: no test)
  const double* chk_c2;
  const double* chk_F;
  int nsp;              // doubles per norm slot (>= 1)
  double* part;         // expm: per-block (max |B| row sum, max |F| row sum), reduced by k_part_reduce
  // optional extra diagonal: A + xdscale * diag(xdiag) (xdscale from rec when set)
  const double* xdiag;
  double2 xdscale;
  // device step record: partition, a, mu, tol, dscale read here (launch = term j of stage i)
  const ExpmSec* rec;
  int stage, term;
};

__device__ __forceinline__ int64_t ld_idx(const void* p, int idx64, int64_t i) {
  return idx64 ? ((const int64_t*)p)[i] : (int64_t)((const int32_t*)p)[i];
}

// G lanes per row.  EPI: fused expm epilogue.
// FT: element type of F (the accumulated Taylor sum) in the expm epilogue;
// FP64 F with FP32 terms is the mixed-precision Taylor loop.
// Expm launch prologue shared by the SpMV kernels: every control value is
// read in one round (no load waits on another).  Returns false when the launch
// has nothing to do (term outside the record's partition, series finished,
// or the truncation test of the previous term passes now); otherwise sets
// the term's coefficient a_t, shift mu_t and extra-diagonal factor xs.
__device__ __forceinline__ bool expm_term_prologue(const CsrArgs& A, double2& a_t, double2& mu_t, double2& xs) {
  const int done = A.ctl->done;
  double tol = A.ctl->tol;
  ExpmSec E{};
  if (A.rec) E = *A.rec;
  double c1 = 0.0, c2 = 0.0, nf = 0.0;
  if (A.chk_c1) { c1 = slot_max(A.chk_c1, A.nsp); c2 = slot_max(A.chk_c2, A.nsp); nf = slot_max(A.chk_F, A.nsp); }
  if (A.rec) {
    if (A.stage >= E.s || A.term >= E.m) return false;
    const double cf = 1.0 / ((double)E.s * (double)(A.term + 1));
    a_t = make_double2(E.a.x * cf, E.a.y * cf);
    mu_t = E.mu;
    xs = E.dscale;
    tol = E.tol;
  }
  if (done) return false;
  // truncation test of the previous term (SCIPY:298-300)
  if (A.chk_c1 && c1 + c2 <= tol * nf) {
    if (blockIdx.x == 0 && threadIdx.x == 0) A.ctl->done = 1;
    return false;
  }
  return true;
}

// Output of one element of an expm term or a plain product:
// res = a_t (acc - (mu - xs d) x_r) (+ y);  with EPI F += rounded res and
// the row sums of |B|, |F| accumulate into curB, curF.
template <typename VT, bool EPI, typename FT>
__device__ __forceinline__ void spmv_out(const CsrArgs& A, int64_t o, int64_t row, double2 acc, double2 xr,
                                         double2 a_t, double2 mu_t, double2 xs, double& curB, double& curF) {
  constexpr bool CPLX = VTr<VT>::cplx;
  VT* __restrict__ y = (VT*)A.y;
  double2 mu_r = mu_t;
  if (A.xdiag) {
    // the extra diagonal enters as a row-dependent shift
    const double d = A.xdiag[row];
    mu_r = make_double2(mu_r.x - xs.x * d, mu_r.y - xs.y * d);
  }
  double2 res;
  if constexpr (CPLX) {
    acc = cadd(acc, make_double2(-(mu_r.x * xr.x - mu_r.y * xr.y), -(mu_r.x * xr.y + mu_r.y * xr.x)));
    res = cmul(a_t, acc);
  } else {
    res = make_double2(a_t.x * (acc.x - mu_r.x * xr.x), 0.0);
  }
  if (A.beta) res = cadd(res, VTr<VT>::ld(y + o));
  VTr<VT>::st(y + o, res);
  if constexpr (EPI) {
    FT* F = (FT*)A.F;
    // the stored (rounded) B enters F, as in the reference
    const double2 rs = VTr<VT>::rnd(res);
    const double2 fs = VTr<FT>::rnd(cadd(VTr<FT>::ld(F + o), rs));
    VTr<FT>::st(F + o, fs);
    curB += CPLX ? cabs2(rs) : fabs(rs.x);
    curF += CPLX ? cabs2(fs) : fabs(fs.x);
  }
}

// Block max of the per-thread row-sum maxima into part[blockIdx] (two
// doubles); k_part_reduce folds the partials of a launch into the slots.
__device__ __forceinline__ void block_norm_partials(const CsrArgs& A, double mb, double mf) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mb = fmax(mb, __shfl_xor(mb, o, 64));
    mf = fmax(mf, __shfl_xor(mf, o, 64));
  }
  __shared__ double sw[2][16];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sw[0][w] = mb; sw[1][w] = mf; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) { mb = fmax(mb, sw[0][i]); mf = fmax(mf, sw[1][i]); }
    A.part[2 * (int64_t)blockIdx.x] = mb;
    A.part[2 * (int64_t)blockIdx.x + 1] = mf;
  }
}

template <typename MT, typename VT, int G, bool EPI, typename FT = VT>
__global__ void __launch_bounds__(256) k_csr_spmv(CsrArgs A) {
  double2 a_t = A.a, mu_t = A.mu, xs = A.xdscale;
  if constexpr (EPI) {
    if (!expm_term_prologue(A, a_t, mu_t, xs)) return;
  }
  constexpr bool CPLX = VTr<VT>::cplx;
  const int lane = threadIdx.x % G;
  double rowB = 0.0, rowF = 0.0;  // max over this group's rows of the row sums of |B|, |F|
  // grid-stride over rows (the expm form runs a capped grid so that the norm
  // atomics stay few); the G lanes of a row always iterate together.  Each
  // XCD takes runs of consecutive blocks (xcd_chunked, as the batched H.v
  // kernel does), so its rows' column gathers share more L2 lines: the
  // L = 34 Heisenberg CSR (17.2 M rows) 2.06 -> 1.90 ms, the XXZ L = 28 CSR
  // (40.1 M rows) unchanged.
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) / G;
  const int64_t bxi = xcd_chunked(blockIdx.x, gridDim.x, gridDim.x >= 16384 ? 256 : 64);
  for (int64_t row = (bxi * blockDim.x + threadIdx.x) / G; row < A.Ns; row += stride) {
    const int64_t k0 = ld_idx(A.indptr, A.idx64, row), k1 = ld_idx(A.indptr, A.idx64, row + 1);
    const MT* __restrict__ val = (const MT*)A.data;
    const VT* __restrict__ x = (const VT*)A.x;
    double curB = 0.0, curF = 0.0;
    for (int64_t j = 0; j < A.nvec; ++j) {
      double sr = 0.0, si = 0.0;
      for (int64_t k = k0 + lane; k < k1; k += G) {
        const int64_t c = ld_idx(A.indices, A.idx64, k);
        const double2 m = MTr<MT>::ld(val + k);
        const double2 v = VTr<VT>::ld(x + c * A.nvec + j);
        if constexpr (CPLX) { sr += m.x * v.x - m.y * v.y; si += m.x * v.y + m.y * v.x; }
        else sr += m.x * v.x;
      }
      sr = grp_sum<G>(sr);
      if constexpr (CPLX) si = grp_sum<G>(si);
      if (lane == 0) {
        const int64_t o = row * A.nvec + j;
        spmv_out<VT, EPI, FT>(A, o, row, make_double2(sr, si), VTr<VT>::ld(x + o), a_t, mu_t, xs, curB, curF);
      }
    }
    rowB = fmax(rowB, curB);
    rowF = fmax(rowF, curF);
  }
  if constexpr (EPI) block_norm_partials(A, rowB, rowF);
}
