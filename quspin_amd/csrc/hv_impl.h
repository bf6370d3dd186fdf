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
  I rep = s2;
  int code = 0, cnt = 1, sign = 1;
  if constexpr (M != 0) {
    OrbitRes R;
    rep = orbit_max<I, M, FERM>(A.C, s2, R);
    code = R.code;
    cnt = R.cnt;
    sign = R.sign;
  }
  const int64_t c = lookup_index<I, LK>(A, rep);
  if (c < 0) return;
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

#define QSX_HV_MSWITCH(nm, BODY)                               \
  switch (nm) {                                                \
    case 0: { constexpr int MM = 0; BODY; } break;             \
    case 1: { constexpr int MM = 1; BODY; } break;             \
    case 2: { constexpr int MM = 2; BODY; } break;             \
    case 3: { constexpr int MM = 3; BODY; } break;             \
    default: { constexpr int MM = -1; BODY; } break;           \
  }

// Dispatch over (maps, fermion sign, lookup, NV).
template <typename I, typename VT>
int launch_hv_rows(const HvArgs& A, int nmaps, bool fermion_maps, int lookup, hipStream_t st) {
  const int blk = 256;
  const dim3 grid((unsigned)qsx_grid(A.Ns, blk));
  const bool nv1 = (A.nvec == 1);
  if (nmaps == 0) {
    if (lookup == LOOKUP_RANK) {
      if (nv1) hipLaunchKernelGGL((k_hv_rows<I, VT, 0, false, LOOKUP_RANK, 1>), grid, dim3(blk), 0, st, A);
      else hipLaunchKernelGGL((k_hv_rows<I, VT, 0, false, LOOKUP_RANK, 0>), grid, dim3(blk), 0, st, A);
    } else {
      if (nv1) hipLaunchKernelGGL((k_hv_rows<I, VT, 0, false, LOOKUP_SEARCH, 1>), grid, dim3(blk), 0, st, A);
      else hipLaunchKernelGGL((k_hv_rows<I, VT, 0, false, LOOKUP_SEARCH, 0>), grid, dim3(blk), 0, st, A);
    }
  } else if (fermion_maps) {
    QSX_HV_MSWITCH(nmaps, {
      if (nv1) hipLaunchKernelGGL((k_hv_rows<I, VT, MM, true, LOOKUP_SEARCH, 1>), grid, dim3(blk), 0, st, A);
      else hipLaunchKernelGGL((k_hv_rows<I, VT, MM, true, LOOKUP_SEARCH, 0>), grid, dim3(blk), 0, st, A);
    });
  } else {
    QSX_HV_MSWITCH(nmaps, {
      if (nv1) hipLaunchKernelGGL((k_hv_rows<I, VT, MM, false, LOOKUP_SEARCH, 1>), grid, dim3(blk), 0, st, A);
      else hipLaunchKernelGGL((k_hv_rows<I, VT, MM, false, LOOKUP_SEARCH, 0>), grid, dim3(blk), 0, st, A);
    });
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    qsx_set_error(std::string("hv launch failed: ") + hipGetErrorString(e));
    return QSX_ERR_HIP;
  }
  return QSX_OK;
}

// per-dtype entry points (defined in hv_<dtype>.hip)
int qsx_hv_launch_f32(const HvArgs& A, int N, int nmaps, bool fm, int lookup, hipStream_t st);
int qsx_hv_launch_f64(const HvArgs& A, int N, int nmaps, bool fm, int lookup, hipStream_t st);
int qsx_hv_launch_c64(const HvArgs& A, int N, int nmaps, bool fm, int lookup, hipStream_t st);
int qsx_hv_launch_c128(const HvArgs& A, int N, int nmaps, bool fm, int lookup, hipStream_t st);

#define QSX_DEFINE_HV_LAUNCH(NAME, VT)                                                      \
  int NAME(const HvArgs& A, int N, int nmaps, bool fm, int lookup, hipStream_t st) {        \
    if (N <= 32) return launch_hv_rows<uint32_t, VT>(A, nmaps, fm, lookup, st);             \
    return launch_hv_rows<uint64_t, VT>(A, nmaps, fm, lookup, st);                          \
  }
