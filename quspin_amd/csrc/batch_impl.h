// Batched H.V over many symmetry sectors and many vectors in one launch.
//
// Each sector k has its own basis, normalisations, characters and lookup
// table (an HvArgs built from an H.v context); the vectors of all sectors are
// stacked row-wise into one (total_rows, ld) array, sector k starting at row
// row_off[k].  Work is cut into chunks of consecutive rows of one sector; one
// wavefront takes one chunk and, for each row,
//   1. lane t evaluates term t of the (Hermitian-conjugated) operator list on
//      the row state: column index and reduced matrix element;
//   2. the non-vanishing (column, element) pairs are compacted into LDS;
//   3. every lane owns one vector column and accumulates element * x[col]
//      over the compacted list (each gather is one coalesced 64-lane read of a
//      whole vector row), in FP64.
// The per-row diagonal is summed lane-parallel.  With EPI the launch is one
// Taylor term of the expm loop for every sector (per-sector partition,
// truncation test and done flag; see qsx_expm_multiply_batch).
#pragma once
#include "sparse_impl.h"

struct WaveChunk {
  int32_t sec;
  int32_t n;      // rows in this chunk
  int64_t begin;  // first sector-local row
};

// One sector's CSR matrix (the CSR form of a batch).
struct CsrSec {
  const void* indptr;       // int32 or int64 (idx64)
  const int32_t* indices;
  const void* data;         // same element type as the vectors
  int64_t Ns;
  int idx64;
  int pad;
};

struct BatchArgs {
  const HvArgs* sec;        // per sector (v_in/v_out unused); matrix-free form
  const CsrSec* csr;        // per sector; CSR form (sec unused)
  const int64_t* row_off;   // first stacked row of each sector
  const WaveChunk* chunk;
  int64_t nchunk;
  int xcd_blocks;           // blocks per XCD slice (grid = 8 * xcd_blocks), or 0: dispatch order
  int ncol;                 // columns handled by this launch (<= 64)
  int64_t ld;               // row stride of x, y, F (elements)
  int64_t col0;             // first column handled by this launch
  const void* x;
  void* y;
  double2 alpha;            // plain apply: y = alpha H x (+ y if beta)
  int beta;
  const double* xdiag;      // optional extra diagonal over the stacked rows: H + xdscale * diag(xdiag)
  double2 xdscale;          // its factor in a plain apply (the expm takes ExpmSec::dscale)
  int* fail;                // tail check: per sector, set when a stage ended unconverged
  int stage, term;          // EPI: Taylor stage i and term j of this launch
  void* F;
  const ExpmSec* es;
  int* done;                // per sector
  double* nB;               // per sector norm slots [k][i][j], stride nslot
  double* nF;
  int64_t nslot;            // s_max * (m_max + 1)
  int mstride;              // m_max + 1
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Order the LDS writes of one wavefront before its following reads (and the
// reads before the next round's writes).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// xcd_blocks > 0: block b of the XCD-interleaved dispatch (block b runs on
// XCD b % 8) is given the chunk range of its XCD's slice, so neighbouring
// rows share an L2.
// xcd_blocks == 0, WIN (elements of at most 8 bytes): each XCD takes runs of
// 64 consecutive blocks (256 on grids of 16,384 blocks or more) inside a
// window all XCDs share (xcd_chunked), so all XCDs work inside one sector's
// vectors at a time and each XCD's rows are neighbours; without WIN, chunks
// in dispatch order.  Measured on the 26 L = 26 Heisenberg sectors x 64
// vectors (CSR form, 650 k blocks): complex64 H.V 9.45 -> 9.12 ms and expm
// 1.87-2.00 -> 1.57 s.  complex128 H.V went 21.3 -> 16.1 ms (16.98 ms with
// xcd_blocks), but its Taylor-term kernel lost (expm 3.70 s against 3.12 s
// with xcd_blocks, which the tuner then no longer picked: it times the plain
// H.V), so complex128 keeps dispatch order.
// The result is wave-uniform and is marked so (readfirstlane): chunk data,
// sector descriptors and row offsets then live in scalar registers and the
// gathers address vector rows as scalar base + lane offset.
template <bool WIN>
__device__ __forceinline__ int64_t wave_chunk_index(const BatchArgs& B) {
  const int b = blockIdx.x;
  int64_t slice_block = B.xcd_blocks ? (int64_t)(b & 7) * B.xcd_blocks + (b >> 3) : (int64_t)b;
  if (WIN && !B.xcd_blocks) slice_block = xcd_chunked(b, gridDim.x, gridDim.x >= 16384 ? 256 : 64);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  return slice_block * (blockDim.x >> 6) + w;
}

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return ((int64_t)hi << 32) | (int64_t)(unsigned)lo;
}

template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  return (T*)uniform64((int64_t)(uintptr_t)p);
}

__device__ __forceinline__ WaveChunk uniform_chunk(WaveChunk c) {
  WaveChunk u;
  u.sec = __builtin_amdgcn_readfirstlane(c.sec);
  u.n = __builtin_amdgcn_readfirstlane(c.n);
  u.begin = uniform64(c.begin);
  return u;
}

template <typename VT>
__device__ __forceinline__ double vabs(double2 v) {
  return VTr<VT>::cplx ? sqrt(v.x * v.x + v.y * v.y) : fabs(v.x);
}

// Wave-parallel diagonal of row state s: lanes split the diagonal terms.
template <typename I, bool CPLX>
__device__ __forceinline__ double2 wave_diag(const HvArgs& A, int64_t r, I s) {
  double2 d = make_double2(0.0, 0.0);
  for (int t = lane_id(); t < A.n_diag; t += 64) {
    const TermDev* T = A.dg + t;
    I s2 = s;
    double mag;
    int ip;
    if (!apply_term<I>(T, A.C->fermion != 0, s2, mag, ip)) continue;
    d = cadd(d, cmul(make_double2(T->jr * mag, T->ji * mag), ipow_c(ip)));
  }
  d.x = wave_sum(d.x);
  if constexpr (CPLX) d.y = wave_sum(d.y);
  if (A.diag_conj) d = cconj(d);
  if (A.diag_arr) {
    if (A.diag_arr_cplx) d = cadd(d, ((const double2*)A.diag_arr)[r]);
    else d.x += ((const double*)A.diag_arr)[r];
  }
  return d;
}

template <typename VT> __device__ __forceinline__ VT vzero();
template <> __device__ __forceinline__ float vzero<float>() { return 0.f; }
template <> __device__ __forceinline__ double vzero<double>() { return 0.0; }
template <> __device__ __forceinline__ float2 vzero<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ __forceinline__ double2 vzero<double2>() { return make_double2(0.0, 0.0); }

template <typename VT> __device__ __forceinline__ double2 to_d2(VT v);
template <> __device__ __forceinline__ double2 to_d2<float>(float v) { return make_double2((double)v, 0.0); }
template <> __device__ __forceinline__ double2 to_d2<double>(double v) { return make_double2(v, 0.0); }
template <> __device__ __forceinline__ double2 to_d2<float2>(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <> __device__ __forceinline__ double2 to_d2<double2>(double2 v) { return v; }

// EPI prologue for sector k: false when this launch's Taylor term does not
// apply to the sector (outside its partition, finished, or truncated now);
// otherwise the term's coefficient a/(s (j+1)) and shift -coef*mu.
__device__ __forceinline__ bool taylor_term_setup(const BatchArgs& B, int k, int lane, double2& coef_s,
                                                  double2& shift, double2& dsc) {
  const ExpmSec E = B.es[k];
  dsc = E.dscale;
  const int i = B.stage, j = B.term;
  if (i >= E.s || j >= E.m) return false;
  if (B.done[k]) return false;
  if (j >= 1) {
    const double* nb = B.nB + k * B.nslot + (int64_t)i * B.mstride;
    const double* nf = B.nF + k * B.nslot + (int64_t)i * B.mstride;
    // truncation test of the previous term (scipy _expm_multiply.py:298-300)
    if (nb[j - 1] + nb[j] <= E.tol * nf[j]) {
      if (lane == 0) B.done[k] = 1;
      return false;
    }
  }
  const double cf = 1.0 / ((double)E.s * (double)(j + 1));
  coef_s = make_double2(E.a.x * cf, E.a.y * cf);
  shift = make_double2(-(coef_s.x * E.mu.x - coef_s.y * E.mu.y), -(coef_s.x * E.mu.y + coef_s.y * E.mu.x));
  return true;
}

__device__ __forceinline__ void taylor_norms_out(const BatchArgs& B, int k, double maxB, double maxF) {
  atomic_max_nonneg(B.nB + k * B.nslot + (int64_t)B.stage * B.mstride + B.term + 1, maxB);
  atomic_max_nonneg(B.nF + k * B.nslot + (int64_t)B.stage * B.mstride + B.term + 1, maxF);
}

// Output of one row: y = coef*acc + shift*x_row (+ y if beta); with EPI the
// rounded y also enters F and the row sums of |y| and |F| feed the norms.
template <typename VT, bool EPI>
__device__ __forceinline__ void row_out(const BatchArgs& B, int64_t row, int lane, bool col_ok, double2 acc,
                                        double2 xr, double2 coef_s, double2 shift, VT* __restrict__ Y,
                                        VT* __restrict__ FF, double& maxB, double& maxF) {
  constexpr bool CPLX = VTr<VT>::cplx;
  double2 res;
  if constexpr (CPLX) res = cadd(cmul(coef_s, acc), cmul(shift, xr));
  else res = make_double2(coef_s.x * acc.x + shift.x * xr.x, 0.0);
  if constexpr (!EPI) {
    if (col_ok) {
      if (B.beta) res = cadd(res, VTr<VT>::ld(Y + row * B.ld + lane));
      VTr<VT>::st(Y + row * B.ld + lane, res);
    }
  } else {
    double rb = 0.0, rf = 0.0;
    if (col_ok) {
      VTr<VT>::st(Y + row * B.ld + lane, res);
      const double2 rs = VTr<VT>::rnd(res);  // the stored (rounded) term enters F
      const double2 f = VTr<VT>::rnd(cadd(VTr<VT>::ld(FF + row * B.ld + lane), rs));
      VTr<VT>::st(FF + row * B.ld + lane, f);
      rb = vabs<VT>(rs);
      rf = vabs<VT>(f);
    }
    maxB = fmax(maxB, wave_sum(rb));
    maxF = fmax(maxF, wave_sum(rf));
  }
}

template <typename I, typename VT, int M, bool FERM, int LK, bool EPI>
__global__ void __launch_bounds__(256) k_hv_wave(BatchArgs B) {
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  __shared__ int32_t s_col[4][64];
  __shared__ double2 s_cf[4][64];
  const int64_t ci = wave_chunk_index<(sizeof(VT) <= 8)>(B);
  const WaveChunk ch = uniform_chunk(ci < B.nchunk ? B.chunk[ci] : WaveChunk{0, 0, 0});
  {
    // the binomial table is the same for every sector of a batch
    const HvArgs& A0 = B.sec[ch.sec];
    stage_binom<LK>(A0.R, sb);
  }
  __syncthreads();
  if (ch.n == 0) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const int k = ch.sec;
  const HvArgs& A = B.sec[k];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const bool col_ok = lane < B.ncol;
  const int lc = col_ok ? lane : 0;

  double2 coef_s = B.alpha, shift = make_double2(0.0, 0.0), dsc = B.xdscale;
  if constexpr (EPI) {
    if (!taylor_term_setup(B, k, lane, coef_s, shift, dsc)) return;
  }
  const double* __restrict__ xdiag = B.xdiag;

  const int64_t base = uniform64(B.row_off[k]);
  const VT* __restrict__ X = (const VT*)B.x + B.col0;
  VT* __restrict__ Y = (VT*)B.y + B.col0;
  VT* __restrict__ FF = EPI ? (VT*)B.F + B.col0 : nullptr;
  double maxB = 0.0, maxF = 0.0;

  for (int rr_i = 0; rr_i < ch.n; ++rr_i) {
    const int64_t r = ch.begin + rr_i;
    const I s = ((const I*)A.basis)[r];
    double rsq = 1.0;
    if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
    const RowRank rrk = row_rank_if<I, LK>(A, sb, s);
    const double2 dg = wave_diag<I, CPLX>(A, r, s);
    const int64_t row = base + r;
    const double2 xr = VTr<VT>::ld(X + row * B.ld + lc);
    // diagonal part (alpha applies to the whole operator; shift is the -coef*mu of expm)
    double2 acc;
    if constexpr (CPLX) acc = cmul(dg, xr);
    else acc = make_double2(dg.x * xr.x, 0.0);
    for (int t0 = 0; t0 < A.n_off; t0 += 64) {
      const int t = t0 + lane;
      int32_t c = -1;
      double2 cf = make_double2(0.0, 0.0);
      if (t < A.n_off) {
        term_element<I, M, FERM, LK, CPLX>(A, sb, rrk, A.off + t, s, rsq, A.off_conj, [&](int64_t cc, double2 v) {
          c = (int32_t)cc;
          cf = v;
        });
      }
      const uint64_t act = __ballot(c >= 0);
      const int slot = __popcll(act & ((1ull << lane) - 1ull));
      if (c >= 0) { s_col[w][slot] = c; s_cf[w][slot] = cf; }
      wave_lds_sync();
      const int na = __popcll(act);
      // groups of 8 gathers in flight; slots past na read the own row with
      // weight 0, lanes past ncol read column 0 (not stored)
      for (int q0 = 0; q0 < na; q0 += 8) {
        VT xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const bool ok = q0 + u < na;
          const int64_t cc = ok ? (int64_t)__builtin_amdgcn_readfirstlane(s_col[w][(q0 + u) & 63]) : r;
          xv[u] = (X + (base + cc) * B.ld)[lc];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double2 m = (q0 + u < na) ? s_cf[w][(q0 + u) & 63] : make_double2(0.0, 0.0);
          const double2 xd = to_d2<VT>(xv[u]);
          if constexpr (CPLX) acc = cadd(acc, cmul(m, xd));
          else acc.x += m.x * xd.x;
        }
      }
      wave_lds_sync();
    }
    if (xdiag) {
      const double d = xdiag[row];
      const double2 w = make_double2(dsc.x * d, dsc.y * d);
      if constexpr (CPLX) acc = cadd(acc, cmul(w, xr));
      else acc.x += w.x * xr.x;
    }
    row_out<VT, EPI>(B, row, lane, col_ok, acc, xr, coef_s, shift, Y, FF, maxB, maxF);
  }
  if constexpr (EPI) {
    if (lane == 0) taylor_norms_out(B, k, maxB, maxF);
  }
}

// ---- CSR form ---------------------------------------------------------------

__device__ __forceinline__ int32_t rlane(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float2 rlane(float2 v, int l) { return make_float2(rlane(v.x, l), rlane(v.y, l)); }
__device__ __forceinline__ double2 rlane(double2 v, int l) { return make_double2(rlane(v.x, l), rlane(v.y, l)); }
__device__ __forceinline__ int64_t rlane64(int64_t v, int l) {
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return ((int64_t)hi << 32) | (int64_t)(unsigned)lo;
}

// Wavefront per row over a chunk of rows of one sector's CSR matrix; lanes
// own vector columns.  The chunk's row pointers are read once (lane q holds
// row pointer q), a row's (column, value) pairs are read lane-parallel and
// broadcast with readlane, gathers are issued in groups of GB so that they
// are in flight together, and the next row's pairs are requested before the
// current row's gathers are consumed.  Chunks hold at most 63 rows.
template <typename VT, bool EPI>
__global__ void __launch_bounds__(256) k_csr_wave(BatchArgs B) {
  const int64_t ci = wave_chunk_index<(sizeof(VT) <= 8)>(B);
  if (ci >= B.nchunk) return;
  const WaveChunk ch = uniform_chunk(B.chunk[ci]);
  const int k = ch.sec;
  const CsrSec S = B.csr[k];
  const int lane = lane_id();
  const bool col_ok = lane < B.ncol;
  constexpr bool CPLX = VTr<VT>::cplx;
  constexpr int GB = sizeof(VT) >= 16 ? 8 : 16;

  double2 coef_s = B.alpha, shift = make_double2(0.0, 0.0), dsc = B.xdscale;
  if constexpr (EPI) {
    if (!taylor_term_setup(B, k, lane, coef_s, shift, dsc)) return;
  }
  const double* __restrict__ xdiag = B.xdiag;
  const int64_t base = uniform64(B.row_off[k]);
  const VT* __restrict__ X = (const VT*)B.x + B.col0;
  VT* __restrict__ Y = (VT*)B.y + B.col0;
  VT* __restrict__ FF = EPI ? (VT*)B.F + B.col0 : nullptr;
  const int32_t* __restrict__ ind = uniform_ptr(S.indices);
  const VT* __restrict__ val = uniform_ptr((const VT*)S.data);
  double maxB = 0.0, maxF = 0.0;

  const int64_t rp = (lane <= ch.n) ? ld_idx(uniform_ptr(S.indptr), __builtin_amdgcn_readfirstlane(S.idx64),
                                             ch.begin + lane) : 0;
  int64_t k0 = rlane64(rp, 0), k1 = rlane64(rp, 1);
  int32_t c_nx = -1;
  VT v_nx = vzero<VT>();
  if (k0 + lane < k1) { c_nx = ind[k0 + lane]; v_nx = val[k0 + lane]; }

  const int lc = col_ok ? lane : 0;
  for (int q = 0; q < ch.n; ++q) {
    const int64_t r_local = ch.begin + q;
    const int64_t row = base + r_local;
    const double2 xr = VTr<VT>::ld(X + row * B.ld + lc);
    double2 acc = make_double2(0.0, 0.0);
    int32_t c_cur = c_nx;
    VT v_cur = v_nx;
    const int64_t kb = k0, ke = k1;
    if (q + 1 < ch.n) { k0 = k1; k1 = rlane64(rp, q + 2); }
    for (int64_t kk = kb; kk < ke; kk += 64) {
      if (kk != kb) {
        c_cur = -1;
        v_cur = vzero<VT>();
        if (kk + lane < ke) { c_cur = ind[kk + lane]; v_cur = val[kk + lane]; }
      }
      const int n = (int)((ke - kk) < 64 ? (ke - kk) : 64);
      if (kk + 64 >= ke && q + 1 < ch.n) {
        // the next row's pairs, requested before this row's gathers are used
        c_nx = -1;
        v_nx = vzero<VT>();
        if (k0 + lane < k1) { c_nx = ind[k0 + lane]; v_nx = val[k0 + lane]; }
      }
      // Branch-free groups: entries past the row's end carry column -1 and
      // value 0 and read the row's own vector row instead; lanes past ncol
      // read column 0 of the gathered row and are not stored.
      for (int q0 = 0; q0 < n; q0 += GB) {
        VT xv[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) {
          const int32_t cu = rlane(c_cur, q0 + u);
          const int64_t cc = cu >= 0 ? (int64_t)cu : r_local;
          xv[u] = (X + (base + cc) * B.ld)[lc];
        }
#pragma unroll
        for (int u = 0; u < GB; ++u) {
          const double2 m = to_d2<VT>(rlane(v_cur, q0 + u));
          const double2 x = to_d2<VT>(xv[u]);
          if constexpr (CPLX) acc = cadd(acc, cmul(m, x));
          else acc.x += m.x * x.x;
        }
      }
    }
    if (xdiag) {
      const double d = xdiag[row];
      const double2 w = make_double2(dsc.x * d, dsc.y * d);
      if constexpr (CPLX) acc = cadd(acc, cmul(w, xr));
      else acc.x += w.x * xr.x;
    }
    row_out<VT, EPI>(B, row, lane, col_ok, acc, xr, coef_s, shift, Y, FF, maxB, maxF);
  }
  if constexpr (EPI) {
    if (lane == 0) taylor_norms_out(B, k, maxB, maxF);
  }
}

// ---- Taylor-loop bookkeeping over the stacked sectors ---------------------

// SCALE = false: B = F and ||F|| -> nB[k][0][0] (loop start).
// SCALE = true : for sectors still in stage i: F *= eta, B = F, done = 0 and
//                ||F|| -> nB[k][i+1][0] when another stage follows.
template <typename VT, bool SCALE>
__global__ void __launch_bounds__(256) k_batch_stage(BatchArgs B) {
  const int64_t ci = wave_chunk_index<(sizeof(VT) <= 8)>(B);
  if (ci >= B.nchunk) return;
  const WaveChunk ch = uniform_chunk(B.chunk[ci]);
  const int k = ch.sec;
  const ExpmSec E = B.es[k];
  const int lane = lane_id();
  const int i = B.stage;
  if (SCALE && i >= E.s) return;
  const double2 eta = expm_eta(E);
  const bool col_ok = lane < B.ncol;
  VT* FF = (VT*)B.F + B.col0;
  VT* Bo = (VT*)B.y + B.col0;
  const int64_t base = uniform64(B.row_off[k]);
  double mx = 0.0;
  for (int q = 0; q < ch.n; ++q) {
    const int64_t row = base + ch.begin + q;
    double a = 0.0;
    if (col_ok) {
      double2 v = VTr<VT>::ld(FF + row * B.ld + lane);
      if constexpr (SCALE) {
        if constexpr (VTr<VT>::cplx) v = cmul(eta, v);
        else v.x *= eta.x;
        v = VTr<VT>::rnd(v);
        VTr<VT>::st(FF + row * B.ld + lane, v);
      }
      VTr<VT>::st(Bo + row * B.ld + lane, v);
      a = vabs<VT>(v);
    }
    mx = fmax(mx, wave_sum(a));
  }
  if (lane == 0) {
    if (SCALE) {
      B.done[k] = 0;
      if (i + 1 < E.s) atomic_max_nonneg(B.nB + k * B.nslot + (int64_t)(i + 1) * B.mstride, mx);
    } else {
      atomic_max_nonneg(B.nB + k * B.nslot, mx);
    }
  }
}

// Launch helpers (instantiated per vector type in batch_<dtype>.hip).
template <typename I, typename VT, bool EPI>
void launch_hv_wave_t(const BatchArgs& B, int lookup, bool fermion, dim3 grid, hipStream_t st) {
  if (lookup == LOOKUP_TABLE) {
    if (fermion) hipLaunchKernelGGL((k_hv_wave<I, VT, 1, true, LOOKUP_TABLE, EPI>), grid, dim3(256), 0, st, B);
    else hipLaunchKernelGGL((k_hv_wave<I, VT, 1, false, LOOKUP_TABLE, EPI>), grid, dim3(256), 0, st, B);
  } else {
    hipLaunchKernelGGL((k_hv_wave<I, VT, 0, false, LOOKUP_RANK, EPI>), grid, dim3(256), 0, st, B);
  }
}

#define QSX_DEFINE_WAVE_LAUNCH(NAME, VT)                                                                  \
  void NAME(const BatchArgs& B, int N, int lookup, bool fermion, int epi, int stage_kind, dim3 grid,       \
            hipStream_t st) {                                                                             \
    if (stage_kind == 1) { hipLaunchKernelGGL((k_batch_stage<VT, false>), grid, dim3(256), 0, st, B); return; } \
    if (stage_kind == 2) { hipLaunchKernelGGL((k_batch_stage<VT, true>), grid, dim3(256), 0, st, B); return; }  \
    if (B.csr) {                                                                                          \
      if (epi) hipLaunchKernelGGL((k_csr_wave<VT, true>), grid, dim3(256), 0, st, B);                     \
      else hipLaunchKernelGGL((k_csr_wave<VT, false>), grid, dim3(256), 0, st, B);                        \
      return;                                                                                             \
    }                                                                                                     \
    if (N <= 32) {                                                                                        \
      if (epi) launch_hv_wave_t<uint32_t, VT, true>(B, lookup, fermion, grid, st);                        \
      else launch_hv_wave_t<uint32_t, VT, false>(B, lookup, fermion, grid, st);                           \
    } else {                                                                                              \
      if (epi) launch_hv_wave_t<uint64_t, VT, true>(B, lookup, fermion, grid, st);                        \
      else launch_hv_wave_t<uint64_t, VT, false>(B, lookup, fermion, grid, st);                           \
    }                                                                                                     \
  }

void qsx_wave_launch_f32(const BatchArgs&, int, int, bool, int, int, dim3, hipStream_t);
void qsx_wave_launch_f64(const BatchArgs&, int, int, bool, int, int, dim3, hipStream_t);
void qsx_wave_launch_c64(const BatchArgs&, int, int, bool, int, int, dim3, hipStream_t);
void qsx_wave_launch_c128(const BatchArgs&, int, int, bool, int, int, dim3, hipStream_t);
