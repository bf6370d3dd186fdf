// Fused, gather-form matrix-free H.v in a symmetry-reduced basis.
//
// Restates what ``basis_general._inplace_Op`` -> ``_core.inplace_op``
// computes (src/quspin/basis/basis_general/base_general.py:354-410): for every
// term J*Op, v_out += a * J * Op v_in with the reduced matrix element
//     H[r', r] = J m sign chi(g) sqrt(n_r' / n_r),   g.s' = r' = rep(s'),
// (m, s') = Op|r>.  The reference scatters one pass per term; here one lane
// owns one output row and GATHERS over all terms in a single launch:
//   row r, Hermitian-conjugate term Op^+ applied to r gives (m', s''), the
//   column c = index(rep(s'')), and H[r, c] = conj((H^+)[c, r]).
// For transposed/conjugated products the host passes either the original or
// the conjugated terms and the conj flag (plain: Op^+ terms, conj; H^T:
// original, no conj; H^*: Op^+, no conj; H^H: original, conj).  Diagonal terms
// are evaluated on the fly.  Accumulation is FP64 for every vector dtype.
#pragma once
#include "qsx_host.h"

template <typename VT> struct VTr;
template <> struct VTr<float> {
  static constexpr bool cplx = false;
  __device__ static double2 ld(const float* p) { return make_double2((double)*p, 0.0); }
  __device__ static void st(float* p, double2 v) { *p = (float)v.x; }
};
template <> struct VTr<double> {
  static constexpr bool cplx = false;
  __device__ static double2 ld(const double* p) { return make_double2(*p, 0.0); }
  __device__ static void st(double* p, double2 v) { *p = v.x; }
};
template <> struct VTr<float2> {
  static constexpr bool cplx = true;
  __device__ static double2 ld(const float2* p) { float2 v = *p; return make_double2((double)v.x, (double)v.y); }
  __device__ static void st(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }
};
template <> struct VTr<double2> {
  static constexpr bool cplx = true;
  __device__ static double2 ld(const double2* p) { return *p; }
  __device__ static void st(double2* p, double2 v) { *p = v; }
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// i^p
__device__ __forceinline__ double2 ipow_c(int p) {
  p &= 3;
  return p == 0 ? make_double2(1, 0) : p == 1 ? make_double2(0, 1) : p == 2 ? make_double2(-1, 0) : make_double2(0, -1);
}

struct HvArgs {
  const CoreDev* C;
  const double2* ph;
  const RankDev* R;
  const TermDev* off;
  const TermDev* dg;
  const void* diag_arr;  // optional per-row diagonal, double2 (complex128) or double
  const void* basis;
  const void* nrm;
  const int64_t* begin;
  const int64_t* end;
  const void* v_in;
  void* v_out;
  int64_t Ns;
  int64_t nvec;
  double2 alpha_off, alpha_diag;
  double2 shift;      // added to the diagonal factor (expm uses -coef*mu)
  const int* skip;    // when non-null and set, the launch does nothing
  int n_off, n_diag;
  int off_conj, diag_conj;
  int diag_arr_cplx;  // 1: double2, 0: double
  int n_bytes;
  int N_p;
  int beta;           // 0: overwrite v_out, 1: accumulate into v_out
  // LOOKUP_TABLE: entry per sector rank = basis index | element code << idx_bits
  // (u32), or index (40 bits) | code << 40 | sign << 63 (u64); all ones = absent
  const void* table;
  int table_u64;
  int idx_bits;
  int sched;          // 0: one lane per row, 1: compacted queue (one vector)
};

__device__ __forceinline__ double load_n(const void* p, int nb, int64_t i) {
  return nb == 1 ? (double)((const uint8_t*)p)[i] : nb == 2 ? (double)((const uint16_t*)p)[i]
                                                              : (double)((const uint32_t*)p)[i];
}

template <typename I, int LK>
__device__ __forceinline__ int64_t lookup_index(const HvArgs& A, I rep) {
  if constexpr (LK == LOOKUP_RANK) {
    return rank_index<I>(A.R, rep);
  } else {
    int64_t lo = 0, hi = A.Ns;
    if (A.N_p > 0) {
      const int64_t p = (int64_t)((uint64_t)rep >> (A.C->N - A.N_p));
      lo = A.begin[p];
      hi = A.end[p];
    }
    return bsearch_desc<I>((const I*)A.basis, lo, hi, rep);
  }
}

// Ascending rank of s inside the particle sector described by R, -1 outside.
template <typename I>
__device__ __forceinline__ int64_t sector_rank(const RankDev* __restrict__ R, I s) {
  if (R->mode == 0) {
    if (popc<I>(s) != R->Np) return -1;
    return comb_rank<I>(R, s);
  } else if (R->mode == 1) {
    const I mlo = (((I)1) << R->L) - 1;
    const I up = s >> R->L, dn = s & mlo;
    if (popc<I>(up) != R->Nup || popc<I>(dn) != R->Ndn) return -1;
    return comb_rank<I>(R, up) * R->cdn + comb_rank<I>(R, dn);
  }
  return (int64_t)s;
}

// Representative of s2, its basis index, element code, stabiliser size and
// fermion sign.  Returns false when the representative is not in the basis.
template <typename I, int M, bool FERM, int LK>
__device__ __forceinline__ bool resolve(const HvArgs& A, I s2, int64_t& c, int& code, int& cnt, int& sign) {
  code = 0; cnt = 1; sign = 1;
  if constexpr (LK == LOOKUP_TABLE) {
    const int64_t rk = sector_rank<I>(A.R, s2);
    if (rk < 0) return false;
    if (A.table_u64) {
      const uint64_t e = ((const uint64_t*)A.table)[rk];
      if (e == ~0ull) return false;
      c = (int64_t)(e & ((1ull << 40) - 1));
      code = (int)((e >> 40) & 0x7FFFFF);
      sign = (e >> 63) ? -1 : 1;
    } else {
      const uint32_t e = ((const uint32_t*)A.table)[rk];
      if (e == 0xFFFFFFFFu) return false;
      c = (int64_t)(e & ((1u << A.idx_bits) - 1u));
      code = (int)(e >> A.idx_bits);
    }
    cnt = (int)load_n(A.nrm, A.n_bytes, c);
    return true;
  } else {
    I rep = s2;
    if constexpr (M != 0) {
      OrbitRes R;
      rep = orbit_max<I, M, FERM>(A.C, s2, R);
      code = R.code;
      cnt = R.cnt;
      sign = R.sign;
    }
    c = lookup_index<I, LK>(A, rep);
    return c >= 0;
  }
}

// Reduced element of term T applied to state s of a row with normalisation
// 1/sqrt(n_r) = rsq.  Calls f(c, coef) when non-zero.  CPLX=false assumes the
// host checked that every factor is real.
template <typename I, int M, bool FERM, int LK, bool CPLX, typename F>
__device__ __forceinline__ void term_element(const HvArgs& A, const TermDev* __restrict__ T, I s, double rsq,
                                             int conj, F&& f) {
  I s2 = s;
  double mag;
  int ip;
  if (!apply_term<I>(T, A.C->fermion != 0, s2, mag, ip)) return;
  int64_t c;
  int code, cnt, sign;
  if (!resolve<I, M, FERM, LK>(A, s2, c, code, cnt, sign)) return;
  double w = mag * (double)sign;
  if constexpr (M != 0) w *= sqrt((double)cnt) * rsq;
  if constexpr (CPLX) {
    double2 coef = cmul(make_double2(T->jr, T->ji), ipow_c(ip));
    coef = make_double2(coef.x * w, coef.y * w);
    if constexpr (M != 0) coef = cmul(coef, A.ph[code]);
    if (conj) coef = cconj(coef);
    f(c, coef);
  } else {
    double v = T->jr * w * ((ip & 2) ? -1.0 : 1.0);
    if constexpr (M != 0) v *= A.ph[code].x;
    f(c, make_double2(v, 0.0));
  }
}

template <typename I, int M, bool FERM, int LK, bool CPLX>
__device__ __forceinline__ double2 row_diag(const HvArgs& A, int64_t r, I s) {
  double2 d = make_double2(0.0, 0.0);
  for (int t = 0; t < A.n_diag; ++t) {
    const TermDev* T = A.dg + t;
    I s2 = s;
    double mag;
    int ip;
    if (!apply_term<I>(T, A.C->fermion != 0, s2, mag, ip)) continue;
    double2 v = cmul(make_double2(T->jr * mag, T->ji * mag), ipow_c(ip));
    d = cadd(d, v);
  }
  if (A.diag_conj) d = cconj(d);
  if (A.diag_arr) {
    if (A.diag_arr_cplx) d = cadd(d, ((const double2*)A.diag_arr)[r]);
    else d.x += ((const double*)A.diag_arr)[r];
  }
  return cadd(cmul(d, A.alpha_diag), A.shift);
}

// One lane per row; NV = 1: register accumulation; NV = 0: any number of
// vectors, accumulated in place in the lane's own output row.
template <typename I, typename VT, int M, bool FERM, int LK, int NV>
__global__ void __launch_bounds__(256) k_hv_rows(HvArgs A) {
  if (A.skip && *A.skip) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const I s = ((const I*)A.basis)[r];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const double2 dsum = row_diag<I, M, FERM, LK, CPLX>(A, r, s);
  if constexpr (NV == 1) {
    double2 acc = make_double2(0.0, 0.0);
    for (int t = 0; t < A.n_off; ++t) {
      term_element<I, M, FERM, LK, CPLX>(A, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
        const double2 x = VTr<VT>::ld(vin + c);
        if constexpr (CPLX) acc = cadd(acc, cmul(coef, x));
        else acc.x += coef.x * x.x;
      });
    }
    double2 res;
    if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc), cmul(dsum, VTr<VT>::ld(vin + r)));
    else res = make_double2(A.alpha_off.x * acc.x + dsum.x * VTr<VT>::ld(vin + r).x, 0.0);
    if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
    VTr<VT>::st(vout + r, res);
  } else {
    const int64_t nv = A.nvec;
    VT* orow = vout + r * nv;
    const VT* irow = vin + r * nv;
    for (int64_t j = 0; j < nv; ++j) {
      double2 res;
      if constexpr (CPLX) res = cmul(dsum, VTr<VT>::ld(irow + j));
      else res = make_double2(dsum.x * VTr<VT>::ld(irow + j).x, 0.0);
      if (A.beta) res = cadd(res, VTr<VT>::ld(orow + j));
      VTr<VT>::st(orow + j, res);
    }
    for (int t = 0; t < A.n_off; ++t) {
      term_element<I, M, FERM, LK, CPLX>(A, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
        double2 cf;
        if constexpr (CPLX) cf = cmul(A.alpha_off, coef);
        else cf = make_double2(A.alpha_off.x * coef.x, 0.0);
        const VT* crow = vin + c * nv;
        for (int64_t j = 0; j < nv; ++j) {
          const double2 x = VTr<VT>::ld(crow + j);
          double2 y = VTr<VT>::ld(orow + j);
          if constexpr (CPLX) y = cadd(y, cmul(cf, x));
          else y.x += cf.x * x.x;
          VTr<VT>::st(orow + j, y);
        }
      });
    }
  }
}

// One vector, compacted work queue.  A block owns 256 rows.  Per chunk of
// QK terms every row appends its non-vanishing (row, term) images s' to an
// LDS queue; the block then resolves the queue densely (representative
// search / rank / table, basis index, character, stabiliser) with every lane
// busy, and each row sums its own entries in term order (deterministic).
constexpr int QBLK = 256;
constexpr int QK = 8;

template <typename I, typename VT, int M, bool FERM, int LK>
__global__ void __launch_bounds__(QBLK) k_hv_queue(HvArgs A) {
  constexpr bool CPLX = VTr<VT>::cplx;
  __shared__ I q_s[QBLK * QK];
  __shared__ double2 q_v[QBLK * QK];
  __shared__ int q_n;
  if (A.skip && *A.skip) return;
  const int tid = threadIdx.x;
  const int64_t r = (int64_t)blockIdx.x * QBLK + tid;
  const bool live = r < A.Ns;
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const I s = live ? ((const I*)A.basis)[r] : (I)0;
  double rsq = 1.0;
  if constexpr (M != 0) rsq = live ? 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r)) : 1.0;
  const bool ferm = A.C->fermion != 0;
  double2 acc = make_double2(0.0, 0.0);
  for (int t0 = 0; t0 < A.n_off; t0 += QK) {
    if (tid == 0) q_n = 0;
    __syncthreads();
    int slot[QK];
    double2 cf[QK];
#pragma unroll
    for (int k = 0; k < QK; ++k) {
      slot[k] = -1;
      cf[k] = make_double2(0.0, 0.0);
      const int t = t0 + k;
      if (live && t < A.n_off) {
        const TermDev* T = A.off + t;
        I s2 = s;
        double mag;
        int ip;
        if (apply_term<I>(T, ferm, s2, mag, ip)) {
          const int sl = atomicAdd(&q_n, 1);
          q_s[sl] = s2;
          slot[k] = sl;
          if constexpr (CPLX) {
            double2 c = cmul(make_double2(T->jr, T->ji), ipow_c(ip));
            c = make_double2(c.x * mag * rsq, c.y * mag * rsq);
            cf[k] = A.off_conj ? cconj(c) : c;
          } else {
            cf[k] = make_double2(T->jr * mag * rsq * ((ip & 2) ? -1.0 : 1.0), 0.0);
          }
        }
      }
    }
    __syncthreads();
    const int n = q_n;
    for (int e = tid; e < n; e += QBLK) {
      int64_t c;
      int code, cnt, sign;
      double2 val = make_double2(0.0, 0.0);
      if (resolve<I, M, FERM, LK>(A, q_s[e], c, code, cnt, sign)) {
        double w = (double)sign;
        if constexpr (M != 0) w *= sqrt((double)cnt);
        const double2 x = VTr<VT>::ld(vin + c);
        if constexpr (CPLX) {
          double2 f = make_double2(w, 0.0);
          if constexpr (M != 0) f = make_double2(w * A.ph[code].x, w * A.ph[code].y);
          if (A.off_conj) f = cconj(f);
          val = cmul(f, x);
        } else {
          double f = w;
          if constexpr (M != 0) f *= A.ph[code].x;
          val = make_double2(f * x.x, 0.0);
        }
      }
      q_v[e] = val;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < QK; ++k) {
      if (slot[k] >= 0) {
        const double2 v = q_v[slot[k]];
        if constexpr (CPLX) acc = cadd(acc, cmul(cf[k], v));
        else acc.x += cf[k].x * v.x;
      }
    }
    __syncthreads();
  }
  if (!live) return;
  const double2 dsum = row_diag<I, M, FERM, LK, CPLX>(A, r, s);
  double2 res;
  if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc), cmul(dsum, VTr<VT>::ld(vin + r)));
  else res = make_double2(A.alpha_off.x * acc.x + dsum.x * VTr<VT>::ld(vin + r).x, 0.0);
  if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
  VTr<VT>::st(vout + r, res);
}

// Scatter form for operator lists that do not commute with the symmetry
// group (the gather identity above needs a symmetric list): one lane per
// source state c applies the ORIGINAL terms and adds J m chi sqrt(n_r/n_c)
// v_in[c] into row r with FP atomics, as the reference's per-term loop does
// (v_out[row] += ME v_in[col], src/quspin/basis/base.py:14-23).
template <typename VT>
__device__ __forceinline__ void atomic_acc(VT* p, double2 v);
template <> __device__ __forceinline__ void atomic_acc<float>(float* p, double2 v) { atomicAdd(p, (float)v.x); }
template <> __device__ __forceinline__ void atomic_acc<double>(double* p, double2 v) { atomicAdd(p, v.x); }
template <> __device__ __forceinline__ void atomic_acc<float2>(float2* p, double2 v) {
  atomicAdd(&p->x, (float)v.x);
  atomicAdd(&p->y, (float)v.y);
}
template <> __device__ __forceinline__ void atomic_acc<double2>(double2* p, double2 v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
}

template <typename I, typename VT, int M, bool FERM, int LK>
__global__ void __launch_bounds__(256) k_hv_scatter(HvArgs A) {
  if (A.skip && *A.skip) return;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const I s = ((const I*)A.basis)[c];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, c));
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* vout = (VT*)A.v_out;
  const int64_t nv = A.nvec;
  for (int t = 0; t < A.n_off; ++t) {
    term_element<I, M, FERM, LK, CPLX>(A, A.off + t, s, rsq, A.off_conj, [&](int64_t r, double2 coef) {
      double2 cf;
      if constexpr (CPLX) cf = cmul(A.alpha_off, coef);
      else cf = make_double2(A.alpha_off.x * coef.x, 0.0);
      for (int64_t j = 0; j < nv; ++j) {
        const double2 x = VTr<VT>::ld(vin + c * nv + j);
        double2 y;
        if constexpr (CPLX) y = cmul(cf, x);
        else y = make_double2(cf.x * x.x, 0.0);
        atomic_acc<VT>(vout + r * nv + j, y);
      }
    });
  }
}

#define QSX_HV_MSWITCH(nm, BODY)                               \
  switch (nm) {                                                \
    case 0: { constexpr int MM = 0; BODY; } break;             \
    case 1: { constexpr int MM = 1; BODY; } break;             \
    case 2: { constexpr int MM = 2; BODY; } break;             \
    case 3: { constexpr int MM = 3; BODY; } break;             \
    default: { constexpr int MM = -1; BODY; } break;           \
  }

// Dispatch over (maps, fermion sign, lookup, schedule, NV).
template <typename I, typename VT, int M, bool FERM, int LK>
void launch_hv_one(const HvArgs& A, hipStream_t st) {
  const int blk = 256;
  const dim3 grid((unsigned)qsx_grid(A.Ns, blk));
  if (A.nvec == 1 && A.sched == 1 && A.n_off > 0)
    hipLaunchKernelGGL((k_hv_queue<I, VT, M, FERM, LK>), grid, dim3(QBLK), 0, st, A);
  else if (A.nvec == 1)
    hipLaunchKernelGGL((k_hv_rows<I, VT, M, FERM, LK, 1>), grid, dim3(blk), 0, st, A);
  else
    hipLaunchKernelGGL((k_hv_rows<I, VT, M, FERM, LK, 0>), grid, dim3(blk), 0, st, A);
}

template <typename I, typename VT>
int launch_hv_rows(const HvArgs& A, int nmaps, bool fermion_maps, int lookup, hipStream_t st) {
  if (lookup == LOOKUP_TABLE) {
    // no orbit search: the map count only matters through n and the characters
    if (fermion_maps) launch_hv_one<I, VT, 1, true, LOOKUP_TABLE>(A, st);
    else launch_hv_one<I, VT, 1, false, LOOKUP_TABLE>(A, st);
  } else if (nmaps == 0) {
    if (lookup == LOOKUP_RANK) launch_hv_one<I, VT, 0, false, LOOKUP_RANK>(A, st);
    else launch_hv_one<I, VT, 0, false, LOOKUP_SEARCH>(A, st);
  } else if (fermion_maps) {
    QSX_HV_MSWITCH(nmaps, (launch_hv_one<I, VT, MM, true, LOOKUP_SEARCH>(A, st)));
  } else {
    QSX_HV_MSWITCH(nmaps, (launch_hv_one<I, VT, MM, false, LOOKUP_SEARCH>(A, st)));
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    qsx_set_error(std::string("hv launch failed: ") + hipGetErrorString(e));
    return QSX_ERR_HIP;
  }
  return QSX_OK;
}

template <typename I, typename VT>
int launch_hv_scatter(const HvArgs& A, int nmaps, bool fermion_maps, int lookup, hipStream_t st) {
  const int blk = 256;
  const dim3 grid((unsigned)qsx_grid(A.Ns, blk));
  if (lookup == LOOKUP_TABLE) {
    if (fermion_maps) hipLaunchKernelGGL((k_hv_scatter<I, VT, 1, true, LOOKUP_TABLE>), grid, dim3(blk), 0, st, A);
    else hipLaunchKernelGGL((k_hv_scatter<I, VT, 1, false, LOOKUP_TABLE>), grid, dim3(blk), 0, st, A);
  } else if (nmaps == 0) {
    if (lookup == LOOKUP_RANK) hipLaunchKernelGGL((k_hv_scatter<I, VT, 0, false, LOOKUP_RANK>), grid, dim3(blk), 0, st, A);
    else hipLaunchKernelGGL((k_hv_scatter<I, VT, 0, false, LOOKUP_SEARCH>), grid, dim3(blk), 0, st, A);
  } else if (fermion_maps) {
    QSX_HV_MSWITCH(nmaps, hipLaunchKernelGGL((k_hv_scatter<I, VT, MM, true, LOOKUP_SEARCH>), grid, dim3(blk), 0, st, A));
  } else {
    QSX_HV_MSWITCH(nmaps, hipLaunchKernelGGL((k_hv_scatter<I, VT, MM, false, LOOKUP_SEARCH>), grid, dim3(blk), 0, st, A));
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    qsx_set_error(std::string("hv scatter launch failed: ") + hipGetErrorString(e));
    return QSX_ERR_HIP;
  }
  return QSX_OK;
}

// per-dtype entry points (defined in hv_<dtype>.hip); scatter = 1 launches the
// scatter kernel (A carries the original terms, no diagonal)

int qsx_hv_launch_f32(const HvArgs& A, int N, int nmaps, bool fm, int lookup, int scatter, hipStream_t st);
int qsx_hv_launch_f64(const HvArgs& A, int N, int nmaps, bool fm, int lookup, int scatter, hipStream_t st);
int qsx_hv_launch_c64(const HvArgs& A, int N, int nmaps, bool fm, int lookup, int scatter, hipStream_t st);
int qsx_hv_launch_c128(const HvArgs& A, int N, int nmaps, bool fm, int lookup, int scatter, hipStream_t st);

#define QSX_DEFINE_HV_LAUNCH(NAME, VT)                                                         \
  int NAME(const HvArgs& A, int N, int nmaps, bool fm, int lookup, int scatter, hipStream_t st) { \
    if (scatter) {                                                                             \
      if (N <= 32) return launch_hv_scatter<uint32_t, VT>(A, nmaps, fm, lookup, st);           \
      return launch_hv_scatter<uint64_t, VT>(A, nmaps, fm, lookup, st);                        \
    }                                                                                          \
    if (N <= 32) return launch_hv_rows<uint32_t, VT>(A, nmaps, fm, lookup, st);                \
    return launch_hv_rows<uint64_t, VT>(A, nmaps, fm, lookup, st);                             \
  }
