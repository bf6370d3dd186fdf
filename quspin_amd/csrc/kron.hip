// Species-factorised H.v for spinful fermions without spatial symmetry.
//
// In the sector (N_up, N_down) of ``spinful_fermion_basis_general`` the
// descending basis is the row-major grid  index = u * N_d + d  of up states u
// and down states d (both descending), and an operator whose off-diagonal
// terms each act on one species is
//     H = H_up (x) 1 + 1 (x) H_down + diag,
// with H_up / H_down the CSR matrices of the same terms in the spinless
// sectors (the Jordan-Wigner string of a same-species pair of operators only
// crosses modes of that species).  One block takes one up row u of psi
// (viewed as an N_u x N_d matrix): the row is staged in LDS once and serves
// every H_down product of the row; the H_up product reads the partner rows
// u' coalesced along d.  Same products as the CSR of H
// (hamiltonian.dot, hamiltonian_core.py:615-789), with per-row HBM traffic
// of one read and one write instead of the CSR stream.
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>

#include "sparse_impl.h"

struct qsx_kron {
  const void* ip_u; const int32_t* ind_u; const void* dat_u; int idx64_u; int64_t Nu; int64_t nnz_u;
  const void* ip_d; const int32_t* ind_d; const void* dat_d; int idx64_d; int64_t Nd; int64_t nnz_d;
  const double* sdiag;  // static diagonal (terms acting on both species), or null
  int mat_dtype;
};

struct KronArgs {
  CsrArgs A;                 // vectors, scales and the expm fields (A.Ns = Nu * Nd)
  const void* ip_u; const int32_t* ind_u; const void* dat_u; int idx64_u; int64_t Nu;
  const void* ip_d; const int32_t* ind_d; const void* dat_d; int idx64_d; int64_t Nd;
  const double* sdiag;
};

constexpr int64_t KRON_MAX_BLOCKS = 65536;
constexpr int64_t KRON_STAGE_BYTES = 64 * 1024;

// STAGE: the up row of psi is staged in LDS (N_d * sizeof(VT) bytes).
template <typename MT, typename VT, bool EPI, typename FT, bool STAGE>
__global__ void __launch_bounds__(256) k_kron(KronArgs K) {
  extern __shared__ unsigned char kron_smem[];
  VT* srow = (VT*)kron_smem;
  const CsrArgs& A = K.A;
  double2 a_t = A.a, mu_t = A.mu, xs = A.xdscale;
  if constexpr (EPI) {
    if (!expm_term_prologue(A, a_t, mu_t, xs)) return;
  }
  constexpr bool CPLX = VTr<VT>::cplx;
  const VT* __restrict__ X = (const VT*)A.x;
  const MT* __restrict__ du = (const MT*)K.dat_u;
  const MT* __restrict__ dd = (const MT*)K.dat_d;
  const int64_t Nd = K.Nd;
  double maxB = 0.0, maxF = 0.0;
  for (int64_t u = blockIdx.x; u < K.Nu; u += gridDim.x) {
    const VT* __restrict__ xu = X + u * Nd;
    if constexpr (STAGE) {
      for (int64_t d = threadIdx.x; d < Nd; d += blockDim.x) srow[d] = xu[d];
      __syncthreads();
    }
    const int64_t ku0 = ld_idx(K.ip_u, K.idx64_u, u), ku1 = ld_idx(K.ip_u, K.idx64_u, u + 1);
    for (int64_t d = threadIdx.x; d < Nd; d += blockDim.x) {
      double sr = 0.0, si = 0.0;
      // 1 (x) H_down: entries of the staged row
      const int64_t kd0 = ld_idx(K.ip_d, K.idx64_d, d), kd1 = ld_idx(K.ip_d, K.idx64_d, d + 1);
      for (int64_t k = kd0; k < kd1; ++k) {
        const int32_t c = K.ind_d[k];
        const double2 m = MTr<MT>::ld(dd + k);
        const double2 v = STAGE ? VTr<VT>::ld(srow + c) : VTr<VT>::ld(xu + c);
        if constexpr (CPLX) { sr += m.x * v.x - m.y * v.y; si += m.x * v.y + m.y * v.x; }
        else sr += m.x * v.x;
      }
      // H_up (x) 1: partner rows u', same column d (coalesced along d)
      for (int64_t k = ku0; k < ku1; ++k) {
        const int64_t c = (int64_t)K.ind_u[k];
        const double2 m = MTr<MT>::ld(du + k);
        const double2 v = VTr<VT>::ld(X + c * Nd + d);
        if constexpr (CPLX) { sr += m.x * v.x - m.y * v.y; si += m.x * v.y + m.y * v.x; }
        else sr += m.x * v.x;
      }
      const int64_t row = u * Nd + d;
      const double2 xr = STAGE ? VTr<VT>::ld(srow + d) : VTr<VT>::ld(xu + d);
      if (K.sdiag) {
        const double g = K.sdiag[row];
        sr += g * xr.x;
        si += g * xr.y;
      }
      double curB = 0.0, curF = 0.0;
      spmv_out<VT, EPI, FT>(A, row, row, make_double2(sr, si), xr, a_t, mu_t, xs, curB, curF);
      maxB = fmax(maxB, curB);
      maxF = fmax(maxF, curF);
    }
    if constexpr (STAGE) __syncthreads();  // the next row overwrites the staged one
  }
  if constexpr (EPI) block_norm_partials(A, maxB, maxF);
}

static int64_t kron_blocks(const qsx_kron* K) {
  const int64_t nb = K->Nu < KRON_MAX_BLOCKS ? K->Nu : KRON_MAX_BLOCKS;
  return nb < 1 ? 1 : nb;
}

template <typename MT, typename VT, bool EPI, typename FT>
static void launch_kron_t(const KronArgs& KA, const qsx_kron* K, hipStream_t st) {
  const dim3 grid((unsigned)kron_blocks(K)), blk(256);
  const int64_t stage = K->Nd * (int64_t)sizeof(VT);
  if (stage <= KRON_STAGE_BYTES)
    hipLaunchKernelGGL((k_kron<MT, VT, EPI, FT, true>), grid, blk, (size_t)stage, st, KA);
  else
    hipLaunchKernelGGL((k_kron<MT, VT, EPI, FT, false>), grid, blk, 0, st, KA);
}

template <bool EPI>
static int launch_kron(const CsrArgs& A, const qsx_kron* K, int vdt, int fdt, hipStream_t st) {
  KronArgs KA;
  std::memset(&KA, 0, sizeof(KA));
  KA.A = A;
  KA.ip_u = K->ip_u; KA.ind_u = K->ind_u; KA.dat_u = K->dat_u; KA.idx64_u = K->idx64_u; KA.Nu = K->Nu;
  KA.ip_d = K->ip_d; KA.ind_d = K->ind_d; KA.dat_d = K->dat_d; KA.idx64_d = K->idx64_d; KA.Nd = K->Nd;
  KA.sdiag = K->sdiag;
  const int mdt = K->mat_dtype;
  if (fdt < 0) fdt = vdt;
  if (fdt == vdt) {
    if (vdt == QSX_F32 && mdt == QSX_F32) launch_kron_t<float, float, EPI, float>(KA, K, st);
    else if (vdt == QSX_F64 && mdt == QSX_F64) launch_kron_t<double, double, EPI, double>(KA, K, st);
    else if (vdt == QSX_C64 && mdt == QSX_F32) launch_kron_t<float, float2, EPI, float2>(KA, K, st);
    else if (vdt == QSX_C64 && mdt == QSX_C64) launch_kron_t<float2, float2, EPI, float2>(KA, K, st);
    else if (vdt == QSX_C128 && mdt == QSX_F64) launch_kron_t<double, double2, EPI, double2>(KA, K, st);
    else if (vdt == QSX_C128 && mdt == QSX_C128) launch_kron_t<double2, double2, EPI, double2>(KA, K, st);
    else { qsx_set_error("unsupported (factor, vector) dtype pair"); return QSX_ERR_ARG; }
  } else if constexpr (EPI) {
    if (vdt == QSX_C64 && fdt == QSX_C128 && mdt == QSX_F32) launch_kron_t<float, float2, EPI, double2>(KA, K, st);
    else if (vdt == QSX_C64 && fdt == QSX_C128 && mdt == QSX_C64) launch_kron_t<float2, float2, EPI, double2>(KA, K, st);
    else if (vdt == QSX_F32 && fdt == QSX_F64 && mdt == QSX_F32) launch_kron_t<float, float, EPI, double>(KA, K, st);
    else { qsx_set_error("unsupported mixed-precision (factor, terms, accumulator) dtypes"); return QSX_ERR_ARG; }
  } else {
    qsx_set_error("accumulator dtype differs from the vector dtype outside the expm");
    return QSX_ERR_ARG;
  }
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_kron_create(const void* ip_u, const void* ind_u, const void* dat_u, int idx64_u, int64_t Nu,
                               int64_t nnz_u, const void* ip_d, const void* ind_d, const void* dat_d, int idx64_d,
                               int64_t Nd, int64_t nnz_d, const double* sdiag, int mat_dtype, qsx_kron** out) {
  if (!out || Nu < 1 || Nd < 1 || !ip_u || !ip_d || (nnz_u > 0 && (!ind_u || !dat_u)) ||
      (nnz_d > 0 && (!ind_d || !dat_d)) || mat_dtype < QSX_F32 || mat_dtype > QSX_C128) {
    qsx_set_error("qsx_kron_create: bad arguments");
    return QSX_ERR_ARG;
  }
  auto* K = new qsx_kron();
  K->ip_u = ip_u; K->ind_u = (const int32_t*)ind_u; K->dat_u = dat_u; K->idx64_u = idx64_u; K->Nu = Nu; K->nnz_u = nnz_u;
  K->ip_d = ip_d; K->ind_d = (const int32_t*)ind_d; K->dat_d = dat_d; K->idx64_d = idx64_d; K->Nd = Nd; K->nnz_d = nnz_d;
  K->sdiag = sdiag;
  K->mat_dtype = mat_dtype;
  *out = K;
  return QSX_OK;
}

extern "C" void qsx_graph_cache_clear();

// Captured step sequences hold the factors' addresses: drop them with the
// operator so that a later operator at recycled addresses is captured anew.
extern "C" void qsx_kron_destroy(qsx_kron* K) {
  qsx_graph_cache_clear();
  delete K;
}

// y = a (H_up (x) 1 + 1 (x) H_down + diag(sdiag) + xs diag(xdiag)) x (+ y if accumulate)
extern "C" int qsx_kron_matvec(const qsx_kron* K, const double* xdiag, double xs_re, double xs_im, const void* x,
                               void* y, int vec_dtype, double a_re, double a_im, int accumulate, void* stream) {
  if (!K || !x || !y) { qsx_set_error("qsx_kron_matvec: bad arguments"); return QSX_ERR_ARG; }
  if ((vec_dtype == QSX_F32 || vec_dtype == QSX_F64) && (a_im != 0.0 || xs_im != 0.0)) {
    qsx_set_error("real vectors with a complex scale");
    return QSX_ERR_COMPLEX;
  }
  CsrArgs A;
  std::memset(&A, 0, sizeof(A));
  A.Ns = K->Nu * K->Nd; A.nvec = 1; A.x = x; A.y = y; A.a = make_double2(a_re, a_im); A.beta = accumulate ? 1 : 0;
  A.xdiag = xdiag; A.xdscale = make_double2(xs_re, xs_im);
  return launch_kron<false>(A, K, vec_dtype, -1, (hipStream_t)stream);
}

// ||H - mu I + xs diag(xdiag)||_1 of the factorised operator: column (u, d)
// sums offu[u] + offd[d] + |du[u] + dd[d] + sdiag[i] + xs xdiag[i] - mu|
// (i = u Nd + d) with the
// factors' off-diagonal column sums offu / offd and diagonals du / dd
// (complex) given; max over the columns into out[0].
__global__ void __launch_bounds__(256) k_kron_norm(const double* offu, const double2* diu, int64_t Nu,
                                                   const double* offd, const double2* did, int64_t Nd,
                                                   const double* sdiag, const double* xdiag, double2 xs, double2 mu,
                                                   double* part) {
  double mx = 0.0;
  const int64_t n = Nu * Nd;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = i / Nd, d = i - u * Nd;
    double2 g = cadd(diu[u], did[d]);
    g = make_double2(g.x - mu.x, g.y - mu.y);
    if (sdiag) g.x += sdiag[i];
    if (xdiag) g = make_double2(g.x + xs.x * xdiag[i], g.y + xs.y * xdiag[i]);
    mx = fmax(mx, offu[u] + offd[d] + sqrt(g.x * g.x + g.y * g.y));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; ++i) mx = fmax(mx, sw[i]);
    part[blockIdx.x] = mx;
  }
}

__global__ void k_kron_norm_fin(const double* part, int nb, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double mx = 0.0;
  for (int i = 0; i < nb; ++i) mx = fmax(mx, part[i]);
  out[0] = mx;
}

// work: at least 1024 doubles of device scratch (block maxima)
extern "C" int qsx_kron_1_norm(const qsx_kron* K, const double* offu, const double* diag_u, const double* offd,
                               const double* diag_d, const double* xdiag, double xs_re, double xs_im, double mu_re,
                               double mu_im, double* work, double* out, void* stream) {
  if (!K || !offu || !diag_u || !offd || !diag_d || !work || !out) { qsx_set_error("qsx_kron_1_norm: bad arguments"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  const int nb = 1024;
  hipLaunchKernelGGL(k_kron_norm, dim3(nb), dim3(256), 0, st, offu, (const double2*)diag_u, K->Nu, offd,
                     (const double2*)diag_d, K->Nd, K->sdiag, xdiag, make_double2(xs_re, xs_im), make_double2(mu_re, mu_im),
                     work);
  QSX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_kron_norm_fin, dim3(1), dim3(64), 0, st, work, nb, out);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// Time steps with device records on the factorised operator (the loop of
// qsx_evolve_csr_steps with the factorised product as the term kernel).
int qsx_evolve_steps_with(int64_t Ns, int64_t nvec, const double* xdiag, int nsteps, const void* es, int s_max,
                          int m_max, void* v, void* work, int vec_dtype, int state_dtype, int* fail, void* ctl_work,
                          int64_t ctl_bytes, int64_t part_blocks,
                          int (*term)(const CsrArgs&, const void*, int, int, hipStream_t), const void* term_ctx,
                          const void* const key_ptr[3], int key_int, int use_graph, hipStream_t st);
int qsx_part_reduce_launch(const CsrArgs& A, int64_t nblocks, hipStream_t st);

static int kron_term(const CsrArgs& A, const void* ctx, int vdt, int fdt, hipStream_t st) {
  const qsx_kron* K = (const qsx_kron*)ctx;
  const int rc = launch_kron<true>(A, K, vdt, fdt, st);
  if (rc) return rc;
  return qsx_part_reduce_launch(A, kron_blocks(K), st);
}

extern "C" int qsx_evolve_kron_steps(const qsx_kron* K, const double* xdiag, int nsteps, const void* es, int s_max,
                                     int m_max, void* v, void* work, int vec_dtype, int state_dtype, int* fail,
                                     void* ctl_work, int64_t ctl_bytes, int use_graph, void* stream) {
  if (!K) { qsx_set_error("qsx_evolve_kron_steps: null operator"); return QSX_ERR_ARG; }
  const void* key[3] = {K->dat_u, K->dat_d, K};
  return qsx_evolve_steps_with(K->Nu * K->Nd, 1, xdiag, nsteps, es, s_max, m_max, v, work, vec_dtype, state_dtype,
                               fail, ctl_work, ctl_bytes, kron_blocks(K), kron_term, K, key, 1024 + K->mat_dtype,
                               use_graph, (hipStream_t)stream);
}
