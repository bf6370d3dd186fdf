// Species-factorised H.v for spinful fermions without spatial symmetry.
//
// In the sector (N_up, N_down) of ``spinful_fermion_basis_general`` the
// descending basis is the row-major grid  index = u * N_d + d  of up states u
// and down states d (both descending), and an operator whose off-diagonal
// terms each act on one species is
//     H = H_up (x) 1 + 1 (x) H_down + diag,
// with H_up / H_down the CSR matrices of the same terms in the spinless
// sectors (the Jordan-Wigner string of a same-species pair of operators only
// crosses modes of that species).  Viewing psi as an N_u x N_d matrix X,
// H psi = H_up X + X H_down^T + diag: the down part combines entries of one
// row, the up part entries of one column.  Same products as the CSR of H
// (hamiltonian.dot, hamiltonian_core.py:615-789).
//
// Two forms:
//  * one pass (small sectors): a block per up row stages the row in LDS for
//    the down part and reads the partner rows for the up part; the partner
//    rows come back from L2 only while the whole vector fits there;
//  * two passes (large sectors): a row pass Z = X H_down^T (row staged in
//    LDS, H_down in a sliced-ELL layout with coalesced entries), then a
//    column-strip pass Y = Z + H_up X + diag over strips of 64 columns, the
//    blocks of one XCD walking a contiguous run of strips so that the strip
//    (N_u x 64 elements) stays in that XCD's L2 while its partner rows are
//    read: each element of X comes from HBM about twice per product instead
//    of once per partner row.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "sparse_impl.h"

struct qsx_kron {
  const void* ip_u; const int32_t* ind_u; const void* dat_u; int idx64_u; int64_t Nu; int64_t nnz_u;
  const void* ip_d; const int32_t* ind_d; const void* dat_d; int idx64_d; int64_t Nd; int64_t nnz_d;
  const double* sdiag;  // static diagonal (terms acting on both species), or null
  int mat_dtype;
  // H_down in sliced ELL: chunks of 64 rows; entry j of row d at
  // sell_off[d / 64] + 64 j + d % 64 (padding: column 0, value 0)
  int64_t* sell_off = nullptr;
  int32_t* sell_idx = nullptr;
  void* sell_val = nullptr;
  // row-pass output of the two-pass form (allocated on first use)
  mutable void* z = nullptr;
  mutable size_t z_bytes = 0;
  int mode = 0;  // 0: by size, 1: one pass, 2: two passes
  // id of the captured step sequences holding these fields (renewed when a
  // field they hold changes)
  mutable uint64_t owner = 0;
};

static void kron_changed(const qsx_kron* K) {
  qsx_graph_drop_owner(K->owner);
  K->owner = qsx_owner_new();
}

struct KronArgs {
  CsrArgs A;                 // vectors, scales and the expm fields (A.Ns = Nu * Nd)
  const void* ip_u; const int32_t* ind_u; const void* dat_u; int idx64_u; int64_t Nu;
  const void* ip_d; const int32_t* ind_d; const void* dat_d; int idx64_d; int64_t Nd;
  const double* sdiag;
  const int64_t* sell_off; const int32_t* sell_idx; const void* sell_val;
  void* z;
};

constexpr int64_t KRON_MAX_BLOCKS = 65536;
constexpr int64_t KRON_STAGE_BYTES = 64 * 1024;

__device__ __forceinline__ int64_t kr_uniform64(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// lane t's double (t wave-uniform)
__device__ __forceinline__ double rl_f64(double x, int t) {
  const int64_t b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), t);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), t);
  return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// acc += m v for a matrix entry m and a vector element v (a real matrix
// entry multiplies both parts of a complex element; complex entries only
// meet complex vectors)
template <typename MT, typename VT>
__device__ __forceinline__ void kr_mac(double& sr, double& si, double2 m, double2 v) {
  if constexpr (MTr<MT>::cplx) { sr += m.x * v.x - m.y * v.y; si += m.x * v.y + m.y * v.x; }
  else if constexpr (VTr<VT>::cplx) { sr += m.x * v.x; si += m.x * v.y; }
  else sr += m.x * v.x;
}

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
        kr_mac<MT, VT>(sr, si, m, v);
      }
      // H_up (x) 1: partner rows u', same column d (coalesced along d)
      for (int64_t k = ku0; k < ku1; ++k) {
        const int64_t c = (int64_t)K.ind_u[k];
        const double2 m = MTr<MT>::ld(du + k);
        const double2 v = VTr<VT>::ld(X + c * Nd + d);
        kr_mac<MT, VT>(sr, si, m, v);
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

// Row pass of the two-pass form: Z[u, :] = X[u, :] H_down^T.  A block stages
// R up rows in LDS (grid-stride over row groups; the staging loads are issued
// four per thread before their LDS stores); each wave takes whole 64-row
// chunks of H_down, so the chunk's width and offset are wave-uniform, the
// sliced-ELL entries load coalesced, and each entry serves the R rows.  The
// staged form runs 1024-thread blocks: many waves share one staged group, so
// a CU keeps 16-32 waves in flight while the H_down entries, which every row
// group reads again from L2, are shared by R rows.  With EPI the launch runs
// the term's prologue, so that a finished series skips both passes.
template <typename MT, typename VT, bool EPI, bool STAGE, int R>
__global__ void __launch_bounds__(1024) k_kron_down(KronArgs K) {
  static_assert(STAGE || R == 1, "direct reads take one row at a time");
  extern __shared__ unsigned char kron_smem[];
  VT* srow = (VT*)kron_smem;
  if constexpr (EPI) {
    double2 a_t, mu_t, xs;
    if (!expm_term_prologue(K.A, a_t, mu_t, xs)) return;
  }
  const VT* __restrict__ X = (const VT*)K.A.x;
  VT* __restrict__ Z = (VT*)K.z;
  const MT* __restrict__ sv = (const MT*)K.sell_val;
  const int32_t* __restrict__ sx = K.sell_idx;
  const int64_t Nd = K.Nd, Nu = K.Nu;
  const int lane = threadIdx.x & 63;
  const int64_t w0 = (int64_t)(threadIdx.x >> 6) * 64;
  for (int64_t u0 = (int64_t)blockIdx.x * R; u0 < Nu; u0 += (int64_t)gridDim.x * R) {
    const int nr = (int)(Nu - u0 < R ? Nu - u0 : R);
    const VT* __restrict__ xu = X + u0 * Nd;
    if constexpr (STAGE) {
      constexpr int SU = 4;
      const int64_t n = nr * Nd, bd = blockDim.x;
      int64_t i = threadIdx.x;
      for (; i + (SU - 1) * bd < n; i += SU * bd) {
        VT t[SU];
#pragma unroll
        for (int q = 0; q < SU; ++q) t[q] = xu[i + q * bd];
#pragma unroll
        for (int q = 0; q < SU; ++q) srow[i + q * bd] = t[q];
      }
      for (; i < n; i += bd) srow[i] = xu[i];
      __syncthreads();
    }
    for (int64_t d0 = w0; d0 < Nd; d0 += blockDim.x) {
      const int64_t ch = d0 >> 6;
      const int64_t off = kr_uniform64(K.sell_off[ch]);
      const int64_t wid = (kr_uniform64(K.sell_off[ch + 1]) - off) >> 6;
      const int32_t* __restrict__ ci = sx + off + lane;
      const MT* __restrict__ cv = sv + off + lane;
      double sr[R], si[R];
#pragma unroll
      for (int r = 0; r < R; ++r) { sr[r] = 0.0; si[r] = 0.0; }
      // entries in batches of four: the batch's index and value loads are
      // issued before its gathers (the width is a multiple of nothing, so a
      // batch past the width reads entry 0 with weight 0)
      for (int64_t j = 0; j < wid; j += 4) {
        int32_t c[4];
        double2 m[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t jq = j + q < wid ? j + q : 0;
          c[q] = ci[jq * 64];
          m[q] = MTr<MT>::ld(cv + jq * 64);
          if (j + q >= wid) m[q] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            // rows past nr read a staged row of this group (discarded)
            const int rr = r < nr ? r : 0;
            const double2 v = STAGE ? VTr<VT>::ld(srow + rr * Nd + c[q]) : VTr<VT>::ld(xu + c[q]);
            kr_mac<MT, VT>(sr[r], si[r], m[q], v);
          }
        }
      }
      const int64_t d = d0 + lane;
      if (d < Nd) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r < nr) st_stream<VT>(Z + (u0 + r) * Nd + d, make_double2(sr[r], si[r]));
      }
    }
    if constexpr (STAGE) __syncthreads();
  }
}

constexpr int KRON_UP_ROWS = 4;  // up rows per wave in the column-strip pass

// Column-strip pass: Y = a_t (Z + H_up X - (mu - diag) X) with the Taylor
// epilogue.  Work unit = (strip s of 64 columns, group g of 16 up rows), in
// strip-major order: wave w takes rows u0 .. u0 + 3 (u0 = 16 g + 4 w; the
// H_up rows are wave-uniform, read through the scalar cache) and lane l
// column d = 64 s + l, so every partner-row read is one coalesced segment of
// the strip; the partner reads of the four rows are issued together.
template <typename MT, typename VT, bool EPI, typename FT>
__global__ void __launch_bounds__(256) k_kron_up(KronArgs K, int64_t nunits, int64_t groups,
                                                 const int32_t* __restrict__ iu, const MT* __restrict__ du) {
  constexpr int RW = KRON_UP_ROWS;
  const CsrArgs& A = K.A;
  double2 a_t = A.a, mu_t = A.mu, xs = A.xdscale;
  if constexpr (EPI) {
    if (!expm_term_prologue(A, a_t, mu_t, xs)) return;
  }
  const VT* __restrict__ X = (const VT*)A.x;
  const VT* __restrict__ Z = (const VT*)K.z;
  const int64_t Nd = K.Nd, Nu = K.Nu;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double maxB = 0.0, maxF = 0.0;
  for (int64_t unit = xcd_contiguous(blockIdx.x, gridDim.x); unit < nunits; unit += gridDim.x) {
    const int64_t s = unit / groups;
    const int64_t u0 = kr_uniform64(((unit - s * groups) * 4 + w) * RW);
    if (u0 >= Nu) continue;
    const int nr = (int)(Nu - u0 < RW ? Nu - u0 : RW);
    const int64_t d = s * 64 + lane;
    const bool dok = d < Nd;
    const int64_t ds = dok ? d : Nd - 1;  // lanes past the last column read in bounds, store nothing
    int64_t kb[RW + 1];
#pragma unroll
    for (int r = 0; r <= RW; ++r) kb[r] = kr_uniform64(ld_idx(K.ip_u, K.idx64_u, u0 + (r < nr ? r : nr)));
    int64_t maxdeg = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) maxdeg = kb[r + 1] - kb[r] > maxdeg ? kb[r + 1] - kb[r] : maxdeg;
    double sr[RW], si[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const double2 z = r < nr ? ld_stream<VT>(Z + (u0 + r) * Nd + ds) : make_double2(0.0, 0.0);
      sr[r] = z.x;
      si[r] = z.y;
    }
    const VT* __restrict__ xd = X + ds;
    const int64_t ne = kb[RW] - kb[0];  // the rows' entries are contiguous
    if (ne <= 64) {
      // one lane-parallel load of all entries, broadcast with readlane, so the
      // only round trips of the loop are the partner reads (issued eight at a
      // time: two steps of the four rows, branch-free; a row past its degree
      // reads entry 0 with weight 0)
      // (lanes past the rows' entries read nothing: with no entries at all kb[0]
      // may be the factor's entry count)
      const int64_t kl = kb[0] + lane;
      const int ci = lane < ne ? iu[kl] : 0;
      const double2 ml = lane < ne ? MTr<MT>::ld(du + kl) : make_double2(0.0, 0.0);
      for (int64_t j = 0; j < maxdeg; j += 2) {
        int64_t c[2][RW];
        double2 m[2][RW];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int r = 0; r < RW; ++r) {
            const int64_t k = kb[r] + j + h;
            const bool on = k < kb[r + 1];  // wave-uniform
            const int t = on ? (int)(k - kb[0]) : 0;
            c[h][r] = (int64_t)__builtin_amdgcn_readlane(ci, t);
            m[h][r] = make_double2(on ? rl_f64(ml.x, t) : 0.0, on ? rl_f64(ml.y, t) : 0.0);
          }
        }
        double2 v[2][RW];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < RW; ++r) v[h][r] = VTr<VT>::ld(xd + c[h][r] * Nd);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < RW; ++r) kr_mac<MT, VT>(sr[r], si[r], m[h][r], v[h][r]);
      }
    } else {
      for (int64_t j = 0; j < maxdeg; ++j) {
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const int64_t k = kb[r] + j;
          if (k < kb[r + 1]) {  // wave-uniform
            const double2 v = VTr<VT>::ld(xd + (int64_t)iu[k] * Nd);
            kr_mac<MT, VT>(sr[r], si[r], MTr<MT>::ld(du + k), v);
          }
        }
      }
    }
    if (!dok) continue;
    // epilogue (spmv_out's arithmetic): every row's operands are requested
    // before the first store, so the rows' loads overlap instead of each row
    // waiting behind the previous row's stores (rows past nr load row u0)
    double2 xr[RW], fo[RW];
    double gd[RW], dx[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int64_t row = (u0 + (r < nr ? r : 0)) * Nd + d;
      xr[r] = VTr<VT>::ld(X + row);
      gd[r] = K.sdiag ? ld_stream<double>(K.sdiag + row).x : 0.0;
      dx[r] = A.xdiag ? ld_stream<double>(A.xdiag + row).x : 0.0;
      if constexpr (EPI) fo[r] = ld_stream<FT>((const FT*)A.F + row);
    }
    VT* __restrict__ y = (VT*)A.y;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      if (r >= nr) break;
      const int64_t row = (u0 + r) * Nd + d;
      const double ar = sr[r] + gd[r] * xr[r].x, ai = si[r] + gd[r] * xr[r].y;
      const double2 mu_r = make_double2(mu_t.x - xs.x * dx[r], mu_t.y - xs.y * dx[r]);
      double2 res;
      if constexpr (VTr<VT>::cplx) {
        const double2 acc = make_double2(ar - (mu_r.x * xr[r].x - mu_r.y * xr[r].y),
                                         ai - (mu_r.x * xr[r].y + mu_r.y * xr[r].x));
        res = cmul(a_t, acc);
      } else {
        res = make_double2(a_t.x * (ar - mu_r.x * xr[r].x), 0.0);
      }
      if (A.beta) res = cadd(res, VTr<VT>::ld(y + row));
      st_stream<VT>(y + row, res);
      if constexpr (EPI) {
        // the stored (rounded) B enters F, as in the reference
        const double2 rs = VTr<VT>::rnd(res);
        const double2 fs = VTr<FT>::rnd(cadd(fo[r], rs));
        st_stream<FT>((FT*)A.F + row, fs);
        maxB = fmax(maxB, VTr<VT>::cplx ? cabs2(rs) : fabs(rs.x));
        maxF = fmax(maxF, VTr<VT>::cplx ? cabs2(fs) : fabs(fs.x));
      }
    }
  }
  if constexpr (EPI) block_norm_partials(A, maxB, maxF);
}

static bool kron_two_pass(const qsx_kron* K, int64_t vsize) {
  if (K->mode == 1) return false;
  if (K->mode == 2) return true;
  // one pass while the vector stays L2-resident (8 XCDs x 4 MB)
  return K->Nu * K->Nd * vsize > ((int64_t)16 << 20);
}

static int64_t kron_up_groups(const qsx_kron* K) { return (K->Nu + 4 * KRON_UP_ROWS - 1) / (4 * KRON_UP_ROWS); }
static int64_t kron_up_units(const qsx_kron* K) { return ((K->Nd + 63) / 64) * kron_up_groups(K); }

// blocks of the launch that writes the norm partials
static int64_t kron_blocks(const qsx_kron* K, int64_t vsize) {
  const int64_t n = kron_two_pass(K, vsize) ? kron_up_units(K) : K->Nu;
  const int64_t nb = n < KRON_MAX_BLOCKS ? n : KRON_MAX_BLOCKS;
  return nb < 1 ? 1 : nb;
}

template <typename MT, typename VT, bool EPI, typename FT>
static void launch_kron_t(const KronArgs& KA, const qsx_kron* K, hipStream_t st) {
  const int64_t stage = K->Nd * (int64_t)sizeof(VT);
  const bool staged = stage <= KRON_STAGE_BYTES;
  const dim3 blk(256);
  if (kron_two_pass(K, sizeof(VT))) {
    // staged row pass: 1024-thread blocks with two rows each while they take
    // at most 128 KB of LDS (measured at L=14, (7,7): complex64 337 -> 369
    // and complex128 147 -> 182 driven steps/s against 256-thread blocks of
    // one row; 768 threads within 1 %, three or four rows and one row slower)
    const int R = !staged ? 1 : (2 * stage <= 128 * 1024) ? 2 : 1;
    const int64_t ng = (K->Nu + R - 1) / R;
    const dim3 gd((unsigned)(ng < KRON_MAX_BLOCKS ? ng : KRON_MAX_BLOCKS));
    if (!staged)
      hipLaunchKernelGGL((k_kron_down<MT, VT, EPI, false, 1>), gd, blk, 0, st, KA);
    else if (R == 2)
      hipLaunchKernelGGL((k_kron_down<MT, VT, EPI, true, 2>), gd, dim3(1024), (size_t)(2 * stage), st, KA);
    else
      hipLaunchKernelGGL((k_kron_down<MT, VT, EPI, true, 1>), gd, dim3(1024), (size_t)stage, st, KA);
    hipLaunchKernelGGL((k_kron_up<MT, VT, EPI, FT>), dim3((unsigned)kron_blocks(K, sizeof(VT))), blk, 0, st, KA,
                       kron_up_units(K), kron_up_groups(K), K->ind_u, (const MT*)K->dat_u);
    return;
  }
  const dim3 grid((unsigned)kron_blocks(K, sizeof(VT)));
  if (staged)
    hipLaunchKernelGGL((k_kron<MT, VT, EPI, FT, true>), grid, blk, (size_t)stage, st, KA);
  else
    hipLaunchKernelGGL((k_kron<MT, VT, EPI, FT, false>), grid, blk, 0, st, KA);
}

static int64_t vt_size(int vdt) { return (vdt == QSX_F32) ? 4 : (vdt == QSX_F64 || vdt == QSX_C64) ? 8 : 16; }

// The row-pass buffer of the two-pass form, allocated outside any capture
// (the entry points call this before launching or recording).  A larger
// buffer replaces the old one and drops captured step sequences, which hold
// its address.
static int kron_prepare(const qsx_kron* K, int vdt) {
  const int64_t vs = vt_size(vdt);
  if (!kron_two_pass(K, vs)) return QSX_OK;
  const size_t need = (size_t)(K->Nu * K->Nd * vs);
  if (K->z_bytes >= need) return QSX_OK;
  kron_changed(K);
  if (K->z) {
    QSX_CHECK_HIP(hipFree(K->z));
    K->z = nullptr;
    K->z_bytes = 0;
  }
  QSX_CHECK_HIP(hipMalloc(&K->z, need));
  K->z_bytes = need;
  return QSX_OK;
}

template <bool EPI>
static int launch_kron(const CsrArgs& A, const qsx_kron* K, int vdt, int fdt, hipStream_t st) {
  KronArgs KA;
  std::memset(&KA, 0, sizeof(KA));
  KA.A = A;
  KA.ip_u = K->ip_u; KA.ind_u = K->ind_u; KA.dat_u = K->dat_u; KA.idx64_u = K->idx64_u; KA.Nu = K->Nu;
  KA.ip_d = K->ip_d; KA.ind_d = K->ind_d; KA.dat_d = K->dat_d; KA.idx64_d = K->idx64_d; KA.Nd = K->Nd;
  KA.sdiag = K->sdiag;
  KA.sell_off = K->sell_off; KA.sell_idx = K->sell_idx; KA.sell_val = K->sell_val;
  KA.z = K->z;
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
  // H_down in sliced ELL, built on the host from the device CSR
  const size_t msz = (mat_dtype == QSX_F32) ? 4 : (mat_dtype == QSX_F64 || mat_dtype == QSX_C64) ? 8 : 16;
  std::vector<int64_t> ip((size_t)Nd + 1);
  if (idx64_d) {
    QSX_CHECK_HIP(hipMemcpy(ip.data(), ip_d, sizeof(int64_t) * (Nd + 1), hipMemcpyDeviceToHost));
  } else {
    std::vector<int32_t> ip32((size_t)Nd + 1);
    QSX_CHECK_HIP(hipMemcpy(ip32.data(), ip_d, sizeof(int32_t) * (Nd + 1), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i <= Nd; ++i) ip[i] = ip32[i];
  }
  if (ip[0] != 0 || ip[Nd] != nnz_d) { qsx_set_error("qsx_kron_create: H_down indptr does not span nnz_d"); return QSX_ERR_ARG; }
  std::vector<int32_t> ind((size_t)nnz_d);
  std::vector<unsigned char> val((size_t)nnz_d * msz);
  if (nnz_d) {
    QSX_CHECK_HIP(hipMemcpy(ind.data(), ind_d, sizeof(int32_t) * nnz_d, hipMemcpyDeviceToHost));
    QSX_CHECK_HIP(hipMemcpy(val.data(), dat_d, msz * nnz_d, hipMemcpyDeviceToHost));
  }
  const int64_t nch = (Nd + 63) / 64;
  std::vector<int64_t> off((size_t)nch + 1, 0);
  for (int64_t c = 0; c < nch; ++c) {
    int64_t w = 0;
    for (int64_t d = c * 64; d < Nd && d < (c + 1) * 64; ++d) w = std::max(w, ip[d + 1] - ip[d]);
    off[c + 1] = off[c] + 64 * w;
  }
  const int64_t ne = off[nch] > 0 ? off[nch] : 1;
  std::vector<int32_t> sidx((size_t)ne, 0);
  std::vector<unsigned char> sval((size_t)ne * msz, 0);
  for (int64_t d = 0; d < Nd; ++d) {
    const int64_t c = d / 64, l = d % 64;
    for (int64_t k = ip[d]; k < ip[d + 1]; ++k) {
      const int32_t col = ind[k];
      if (col < 0 || col >= Nd) { qsx_set_error("qsx_kron_create: H_down column out of range"); return QSX_ERR_ARG; }
      const int64_t e = off[c] + 64 * (k - ip[d]) + l;
      sidx[e] = col;
      std::memcpy(&sval[e * msz], &val[k * msz], msz);
    }
  }
  auto* K = new qsx_kron();
  K->ip_u = ip_u; K->ind_u = (const int32_t*)ind_u; K->dat_u = dat_u; K->idx64_u = idx64_u; K->Nu = Nu; K->nnz_u = nnz_u;
  K->ip_d = ip_d; K->ind_d = (const int32_t*)ind_d; K->dat_d = dat_d; K->idx64_d = idx64_d; K->Nd = Nd; K->nnz_d = nnz_d;
  K->sdiag = sdiag;
  K->mat_dtype = mat_dtype;
  hipError_t e1 = hipMalloc(&K->sell_off, sizeof(int64_t) * (nch + 1));
  hipError_t e2 = hipMalloc(&K->sell_idx, sizeof(int32_t) * ne);
  hipError_t e3 = hipMalloc(&K->sell_val, msz * ne);
  if (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess) {
    e1 = hipMemcpy(K->sell_off, off.data(), sizeof(int64_t) * (nch + 1), hipMemcpyHostToDevice);
    e2 = hipMemcpy(K->sell_idx, sidx.data(), sizeof(int32_t) * ne, hipMemcpyHostToDevice);
    e3 = hipMemcpy(K->sell_val, sval.data(), msz * ne, hipMemcpyHostToDevice);
  }
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    if (K->sell_off) (void)hipFree(K->sell_off);
    if (K->sell_idx) (void)hipFree(K->sell_idx);
    if (K->sell_val) (void)hipFree(K->sell_val);
    delete K;
    qsx_set_error("qsx_kron_create: device allocation or copy failed");
    return QSX_ERR_HIP;
  }
  K->owner = qsx_owner_new();
  *out = K;
  return QSX_OK;
}

// 0: one or two passes by size (default), 1: one pass, 2: two passes
extern "C" int qsx_kron_set_mode(qsx_kron* K, int mode) {
  if (!K || mode < 0 || mode > 2) { qsx_set_error("qsx_kron_set_mode: bad arguments"); return QSX_ERR_ARG; }
  K->mode = mode;
  kron_changed(K);
  return QSX_OK;
}

// Captured step sequences hold the factors' addresses: drop them with the
// operator so that a later operator at recycled addresses is captured anew.
extern "C" void qsx_kron_destroy(qsx_kron* K) {
  if (!K) return;
  qsx_graph_drop_owner(K->owner);
  if (K->sell_off) (void)hipFree(K->sell_off);
  if (K->sell_idx) (void)hipFree(K->sell_idx);
  if (K->sell_val) (void)hipFree(K->sell_val);
  if (K->z) (void)hipFree(K->z);
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
  const int rc = kron_prepare(K, vec_dtype);
  if (rc) return rc;
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
  for (int64_t u = blockIdx.x; u < Nu; u += gridDim.x) {
    const double ou = offu[u];
    const double2 gu = make_double2(diu[u].x - mu.x, diu[u].y - mu.y);
    for (int64_t d = threadIdx.x; d < Nd; d += blockDim.x) {
      const int64_t i = u * Nd + d;
      double2 g = cadd(gu, did[d]);
      if (sdiag) g.x += sdiag[i];
      if (xdiag) g = make_double2(g.x + xs.x * xdiag[i], g.y + xs.y * xdiag[i]);
      mx = fmax(mx, ou + offd[d] + sqrt(g.x * g.x + g.y * g.y));
    }
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

__global__ void __launch_bounds__(1024) k_kron_norm_fin(const double* part, int nb, double* out) {
  double mx = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) mx = fmax(mx, part[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  __shared__ double sw[16];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) mx = fmax(mx, sw[i]);
    out[0] = mx;
  }
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
  hipLaunchKernelGGL(k_kron_norm_fin, dim3(1), dim3(1024), 0, st, work, nb, out);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// Time steps with device records on the factorised operator (the loop of
// qsx_evolve_csr_steps with the factorised product as the term kernel).
int qsx_evolve_steps_with(int64_t Ns, int64_t nvec, const double* xdiag, int nsteps, const void* es, int s_max,
                          int m_max, void* v, void* work, int vec_dtype, int state_dtype, int* fail, void* ctl_work,
                          int64_t ctl_bytes, int64_t part_blocks,
                          int (*term)(const CsrArgs&, const void*, int, int, hipStream_t), const void* term_ctx,
                          const StepOpKey& key, int use_graph, hipStream_t st);
int qsx_part_reduce_launch(const CsrArgs& A, int64_t nblocks, hipStream_t st);

static int kron_term(const CsrArgs& A, const void* ctx, int vdt, int fdt, hipStream_t st) {
  const qsx_kron* K = (const qsx_kron*)ctx;
  const int rc = launch_kron<true>(A, K, vdt, fdt, st);
  if (rc) return rc;
  return qsx_part_reduce_launch(A, kron_blocks(K, vt_size(vdt)), st);
}

extern "C" int qsx_evolve_kron_steps(const qsx_kron* K, const double* xdiag, int nsteps, const void* es, int s_max,
                                     int m_max, void* v, void* work, int vec_dtype, int state_dtype, int* fail,
                                     void* ctl_work, int64_t ctl_bytes, int use_graph, void* stream) {
  if (!K) { qsx_set_error("qsx_evolve_kron_steps: null operator"); return QSX_ERR_ARG; }
  const int rc = kron_prepare(K, vec_dtype);
  if (rc) return rc;
  const StepOpKey key{K->owner, {K->dat_u, K->z, K}, 0, 0, 1024 + 16 * K->mode + K->mat_dtype};
  return qsx_evolve_steps_with(K->Nu * K->Nd, 1, xdiag, nsteps, es, s_max, m_max, v, work, vec_dtype, state_dtype,
                               fail, ctl_work, ctl_bytes, kron_blocks(K, vt_size(vec_dtype)), kron_term, K, key,
                               use_graph, (hipStream_t)stream);
}
