// Operators between symmetry sectors and projections between a
// symmetry-reduced basis and the full space (or its particle sector).
//
// The projector of a reduced basis (``get_proj``,
// src/quspin/basis/basis_general/base_general.py:521-582) has one column per
// representative r:
//     P[:, r] = c_r sum_{g in G} chi(g) sign(g, r) |g.r>,   c_r = 1/sqrt(n_r |G|)
// (the coefficient of base_general.py:571-575; chi(g) the character of the
// sector, sign(g, r) the fermion sign of the permutation).  Every kernel
// here walks the orbit of a state with the same visit_orbit as the basis
// construction, so images, characters and signs are those of the H.v.
//
//   Op_shift_sector  v_out += P_2^H Op P_1 v_in   (base_general.py:412-519;
//                    the identity its docstring states and
//                    test/test_Op_shift_sector.py:31-46 checks)
//   get_amp          C = psi_s / psi_r^sym = P[s, r] (base_general.py:1183-1269)
//   project_from     v_full = P v      (project_from_dense, :684-760)
//   project_to       v = P^H v_full    (project_to_dense, :584-666;
//                    test/test_project_to.py:24-29)
//   get_proj         P as CSC arrays   (get_proj, :521-582)
//
// Rows of the full space are the 2^N states in descending order (index
// 2^N - 1 - s), or with pcon the states of the basis' particle sector in
// descending order (index = sector size - 1 - combinadic rank).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <string>

#include "hv_impl.h"

namespace {

bool chars_real(const qsx_core* c) {
  for (const auto& p : c->phase_h)
    if (p.y != 0.0) return false;
  return true;
}

// index of state t in the full space: descending 2^N states, or descending
// states of the recorded particle sector (pcon)
template <typename I>
__device__ __forceinline__ int64_t full_index(const RankDev* __restrict__ R, const int64_t* sb, int pcon, int N, I t) {
  if (pcon) {
    const int64_t rk = sector_rank<I>(R, sb, t);
    return rk < 0 ? (int64_t)-1 : R->Ns - 1 - rk;
  }
  const uint64_t top = (N >= 64) ? ~0ull : ((1ull << N) - 1ull);
  return (int64_t)(top - (uint64_t)t);
}

template <int PCON>
__device__ __forceinline__ void stage_rank(const RankDev* __restrict__ R, int64_t* sb) {
  if constexpr (PCON) stage_binom<LOOKUP_RANK>(R, sb);
}

// c_r = 1/sqrt(n_r |G|)
__device__ __forceinline__ double proj_coef(const void* nrm, int n_bytes, int64_t r, int G, bool has_maps) {
  return has_maps ? 1.0 / sqrt(load_n(nrm, n_bytes, r) * (double)G) : 1.0;
}

// ---------------------------------------------------------------------------
// Op_shift_sector: lane = source row i of basis 1.  For every group element g
// (image u = g.r1, character chi_1(g), sign) the term gives (m, s'); s' is
// resolved in basis 2 (its representative r2 = k.s', factor sign_k chi_2(k)
// sqrt(n_2)).  Summed over g, with c_1 c_2 n_2 = sqrt(n_2 / n_1) / |G|:
//   v_out[r2] += J m chi_1(g) sign_g sign_k chi_2(k) sqrt(n_2/n_1) / |G| v_in[r1]
// ---------------------------------------------------------------------------
template <typename I, int M, bool FERM, typename VT>
__global__ void __launch_bounds__(256) k_shift_sector(HvArgs A2, TermDev T, const CoreDev* __restrict__ C1,
                                                      const double2* __restrict__ ph1, const I* __restrict__ basis1,
                                                      const void* __restrict__ n1, int n1_bytes, int64_t Ns1,
                                                      const VT* __restrict__ vin, VT* vout, int64_t nv, double invG) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Ns1) return;
  const I r1 = basis1[i];
  const double w0 = (M != 0) ? invG / sqrt(load_n(n1, n1_bytes, i)) : 1.0;
  const bool fkind = A2.C->fermion != 0;
  const double2 J = make_double2(T.jr, T.ji);
  auto vis = [&](I u, int code, int sign) -> bool {
    I s2 = u;
    double mag;
    int ip;
    if (!apply_term<I>(&T, fkind, s2, mag, ip)) return true;
    int64_t c;
    double2 fac;
    if (!resolve<I, M, FERM, LOOKUP_SEARCH>(A2, nullptr, s2, c, fac)) return true;
    double2 coef = cmul(J, ipow_c(ip));
    const double w = w0 * mag * (double)sign;
    coef = make_double2(coef.x * w, coef.y * w);
    if constexpr (M != 0) coef = cmul(cmul(coef, ph1[code]), fac);
    for (int64_t j = 0; j < nv; ++j) {
      const double2 x = VTr<VT>::ld(vin + i * nv + j);
      atomic_acc<VT>(vout + c * nv + j, VTr<VT>::cplx ? cmul(coef, x) : make_double2(coef.x * x.x, 0.0));
    }
    return true;
  };
  visit_orbit<I, M, FERM>(C1, r1, vis);
}

// ---------------------------------------------------------------------------
// get_amp: C = P[s, r] for r = rep(s) = k.s: the images h.r = s are
// h = k^{-1} st (st in the stabiliser of r), whose weights sum to
// c_r conj(chi(k)) sign_k n_r = conj(chi(k)) sign_k sqrt(n_r / |G|), with
// n_r = sum_{st} chi(st) sign(st) (0 outside the sector).
// ---------------------------------------------------------------------------
template <typename I, int M, bool FERM, typename OT>
__global__ void __launch_bounds__(256) k_get_amp(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                                 I* states, OT* __restrict__ out, int64_t n, int write_rep,
                                                 double invG) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const I s = states[i];
  double2 amp = make_double2(1.0, 0.0);
  I rep = s;
  if constexpr (M != 0) {
    OrbitRes R;
    rep = orbit_max<I, M, FERM>(C, s, R);
    double nr = 0.0;
    auto vis = [&](I t, int code, int sign) -> bool {
      if (t == rep) nr += sign * ph[code].x;
      return true;
    };
    visit_orbit<I, M, FERM>(C, rep, vis);
    nr = rint(nr);
    const double w = nr > 0.0 ? (double)R.sign * sqrt(nr * invG) : 0.0;
    amp = make_double2(w * ph[R.code].x, -w * ph[R.code].y);
  }
  VTr<OT>::st(out + i, amp);
  if (write_rep) states[i] = rep;
}

// ---------------------------------------------------------------------------
// project_from_dense: lane = reduced row r writes P[:, r] v[r] into its own
// orbit's rows (orbits of distinct representatives are disjoint; images
// repeated by the stabiliser add in the same lane).  v_full starts at zero.
// ---------------------------------------------------------------------------
template <typename I, int M, bool FERM, int PCON, typename VT>
__global__ void __launch_bounds__(256) k_project_from(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                                      const RankDev* __restrict__ R, const I* __restrict__ basis,
                                                      const void* __restrict__ nrm, int n_bytes, int64_t Ns,
                                                      const VT* __restrict__ v, VT* vfull, int64_t nv) {
  __shared__ int64_t sb[PCON ? QSX_SB : 1];
  stage_rank<PCON>(R, sb);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Ns) return;
  const double cr = proj_coef(nrm, n_bytes, r, C->G, M != 0);
  const int N = C->N;
  auto vis = [&](I t, int code, int sign) -> bool {
    const int64_t row = full_index<I>(R, sb, PCON, N, t);
    if (row < 0) return true;
    double2 w = make_double2(cr * sign, 0.0);
    if constexpr (M != 0) w = make_double2(w.x * ph[code].x, w.x * ph[code].y);
    for (int64_t j = 0; j < nv; ++j) {
      const double2 x = VTr<VT>::ld(v + r * nv + j);
      double2 y = VTr<VT>::ld(vfull + row * nv + j);
      if constexpr (VTr<VT>::cplx) y = cadd(y, cmul(w, x));
      else y.x += w.x * x.x;
      VTr<VT>::st(vfull + row * nv + j, y);
    }
    return true;
  };
  visit_orbit<I, M, FERM>(C, basis[r], vis);
}

// project_to_dense: v[r] = (P^H v_full)[r], a gather over the orbit of r.
template <typename I, int M, bool FERM, int PCON, typename VT>
__global__ void __launch_bounds__(256) k_project_to(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                                    const RankDev* __restrict__ R, const I* __restrict__ basis,
                                                    const void* __restrict__ nrm, int n_bytes, int64_t Ns,
                                                    const VT* __restrict__ vfull, VT* __restrict__ v, int64_t nv) {
  __shared__ int64_t sb[PCON ? QSX_SB : 1];
  stage_rank<PCON>(R, sb);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Ns) return;
  const double cr = proj_coef(nrm, n_bytes, r, C->G, M != 0);
  const int N = C->N;
  for (int64_t j = 0; j < nv; ++j) {
    double2 acc = make_double2(0.0, 0.0);
    auto vis = [&](I t, int code, int sign) -> bool {
      const int64_t row = full_index<I>(R, sb, PCON, N, t);
      if (row < 0) return true;
      double2 w = make_double2(cr * sign, 0.0);
      if constexpr (M != 0) w = make_double2(w.x * ph[code].x, -w.x * ph[code].y);  // conj(chi)
      const double2 x = VTr<VT>::ld(vfull + row * nv + j);
      if constexpr (VTr<VT>::cplx) acc = cadd(acc, cmul(w, x));
      else acc.x += w.x * x.x;
      return true;
    };
    visit_orbit<I, M, FERM>(C, basis[r], vis);
    VTr<VT>::st(v + r * nv + j, acc);
  }
}

// get_proj, pass 0: entries per column = distinct orbit images = |G| / |Stab|
// (inside the full space); pass 1: the column's rows ascending, duplicates
// (images repeated by the stabiliser) summed.
template <typename I, int M, bool FERM, int PCON>
__global__ void __launch_bounds__(256) k_get_proj(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                                  const RankDev* __restrict__ R, const I* __restrict__ basis,
                                                  const void* __restrict__ nrm, int n_bytes, int64_t Ns, int pass,
                                                  int64_t* __restrict__ cnt, const int64_t* __restrict__ offs,
                                                  int64_t* __restrict__ rows, double2* __restrict__ vals) {
  __shared__ int64_t sb[PCON ? QSX_SB : 1];
  stage_rank<PCON>(R, sb);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Ns) return;
  const double cr = proj_coef(nrm, n_bytes, r, C->G, M != 0);
  const int N = C->N;
  const I s = basis[r];
  if (pass == 0) {
    OrbitRes O;
    (void)orbit_max<I, M, FERM>(C, s, O);  // O.cnt = |Stab(s)| (s is its orbit's maximum)
    int64_t k = 0;
    auto vis = [&](I t, int, int) -> bool {
      if (full_index<I>(R, sb, PCON, N, t) >= 0) ++k;
      return true;
    };
    visit_orbit<I, M, FERM>(C, s, vis);
    cnt[r] = k / (O.cnt > 0 ? O.cnt : 1);
    return;
  }
  const int64_t o = offs[r], cap = cnt[r];
  int64_t k = 0;
  auto vis = [&](I t, int code, int sign) -> bool {
    const int64_t row = full_index<I>(R, sb, PCON, N, t);
    if (row < 0) return true;
    double2 w = make_double2(cr * sign, 0.0);
    if constexpr (M != 0) w = make_double2(w.x * ph[code].x, w.x * ph[code].y);
    // insert keeping rows ascending; a repeated row adds to its entry
    int64_t p = 0;
    while (p < k && rows[o + p] < row) ++p;
    if (p < k && rows[o + p] == row) {
      vals[o + p] = cadd(vals[o + p], w);
    } else if (k < cap) {
      for (int64_t q = k; q > p; --q) { rows[o + q] = rows[o + q - 1]; vals[o + q] = vals[o + q - 1]; }
      rows[o + p] = row;
      vals[o + p] = w;
      ++k;
    }
    return true;
  };
  visit_orbit<I, M, FERM>(C, s, vis);
}

__global__ void k_proj_write(int64_t Ns, const int64_t* __restrict__ cnt, const int64_t* __restrict__ offs,
                             const int64_t* __restrict__ rows, const double2* __restrict__ vals, void* indptr,
                             void* indices, void* data, int idx_bytes, int dtype) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > Ns) return;
  const int64_t o = (r < Ns) ? offs[r] : offs[Ns - 1] + cnt[Ns - 1];
  if (idx_bytes == 4) ((int32_t*)indptr)[r] = (int32_t)o;
  else ((int64_t*)indptr)[r] = o;
  if (r == Ns) return;
  for (int64_t k = 0; k < cnt[r]; ++k) {
    const int64_t e = o + k;
    if (idx_bytes == 4) ((int32_t*)indices)[e] = (int32_t)rows[e];
    else ((int64_t*)indices)[e] = rows[e];
    const double2 v = vals[e];
    switch (dtype) {
      case QSX_F32: ((float*)data)[e] = (float)v.x; break;
      case QSX_F64: ((double*)data)[e] = v.x; break;
      case QSX_C64: ((float2*)data)[e] = make_float2((float)v.x, (float)v.y); break;
      default: ((double2*)data)[e] = v; break;
    }
  }
}

// (variadic: a launch expands to <<<...>>> with top-level commas)
#define QSX_SEC_MSWITCH(nm, ...)                              \
  switch (nm) {                                               \
    case 0: { constexpr int MM = 0; __VA_ARGS__; } break;     \
    case 1: { constexpr int MM = 1; __VA_ARGS__; } break;     \
    case 2: { constexpr int MM = 2; __VA_ARGS__; } break;     \
    case 3: { constexpr int MM = 3; __VA_ARGS__; } break;     \
    default: { constexpr int MM = -1; __VA_ARGS__; } break;   \
  }

// I (state word), FM (fermion maps) and MM (map count) dispatch
#define QSX_SEC_DISPATCH(c, ...)                                                          \
  do {                                                                                    \
    const bool fm_ = (c)->h.fermion != 0;                                                 \
    if ((c)->h.N <= 32) {                                                                 \
      using I = uint32_t;                                                                 \
      if (fm_) { constexpr bool FM = true; QSX_SEC_MSWITCH((c)->h.nmaps, __VA_ARGS__); }  \
      else { constexpr bool FM = false; QSX_SEC_MSWITCH((c)->h.nmaps, __VA_ARGS__); }     \
    } else {                                                                              \
      using I = uint64_t;                                                                 \
      if (fm_) { constexpr bool FM = true; QSX_SEC_MSWITCH((c)->h.nmaps, __VA_ARGS__); }  \
      else { constexpr bool FM = false; QSX_SEC_MSWITCH((c)->h.nmaps, __VA_ARGS__); }     \
    }                                                                                     \
  } while (0)

int check_pcon(const qsx_core* c, int pcon, int64_t Ns_full) {
  if (pcon) {
    if (!c->sector_ok || c->rank_h.mode > 1) {
      qsx_set_error("pcon=True only works for basis of a single particle number sector.");
      return QSX_ERR_ARG;
    }
    if (Ns_full != c->rank_h.Ns) {
      qsx_set_error("full-space size does not match the particle sector");
      return QSX_ERR_ARG;
    }
  } else {
    if (c->h.N > 62 || Ns_full != ((int64_t)1 << c->h.N)) {
      qsx_set_error("full-space size must be 2^N (N <= 62)");
      return QSX_ERR_ARG;
    }
  }
  return QSX_OK;
}

}  // namespace

extern "C" int qsx_op_shift_sector(qsx_core* core, qsx_core* other_core, const void* v_in, void* v_out, int dtype,
                                   int64_t nvec, const char* opstr, const int32_t* indx, int n_op, double J_re,
                                   double J_im, const void* basis, const void* n, int n_bytes, int64_t Ns,
                                   const int64_t* begin, const int64_t* end, int N_p, const void* other_basis,
                                   const void* other_n, int other_n_bytes, int64_t other_Ns, void* stream) {
  if (!core || !other_core || nvec < 1 || (dtype != QSX_F64 && dtype != QSX_C128)) {
    qsx_set_error("qsx_op_shift_sector: bad arguments (dtype float64 or complex128)");
    return QSX_ERR_ARG;
  }
  if (core->h.N != other_core->h.N || core->h.nmaps != other_core->h.nmaps || core->h.kind != other_core->h.kind ||
      core->h.G != other_core->h.G) {
    qsx_set_error("Op_shift_sector: the two bases must have the same symmetries and site count");
    return QSX_ERR_ARG;
  }
  TermDev T;
  int ny = 0;
  int rc = qsx_compile_term(core, opstr, indx, n_op, J_re, J_im, 0, T, ny);
  if (rc) return rc;
  if (dtype == QSX_F64 && (J_im != 0.0 || (ny & 1) || !chars_real(core) || !chars_real(other_core))) {
    qsx_set_error("real dtype requested for an operator with complex matrix elements");
    return QSX_ERR_COMPLEX;
  }
  if (Ns == 0 || other_Ns == 0) return QSX_OK;
  if (!v_in || !v_out || !basis || !other_basis) { qsx_set_error("qsx_op_shift_sector: null array"); return QSX_ERR_ARG; }
  HvArgs A;
  std::memset(&A, 0, sizeof(A));
  A.C = core->d; A.ph = core->phase; A.R = core->rank_d;
  A.basis = basis; A.nrm = n; A.n_bytes = n_bytes; A.begin = begin; A.end = end; A.N_p = N_p; A.Ns = Ns;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)qsx_grid(other_Ns, 256)), blk(256);
  const double invG = 1.0 / (double)core->h.G;
  QSX_SEC_DISPATCH(core, {
    if (dtype == QSX_F64)
      hipLaunchKernelGGL((k_shift_sector<I, MM, FM, double>), grid, blk, 0, st, A, T, other_core->d,
                         other_core->phase, (const I*)other_basis, other_n, other_n_bytes, other_Ns,
                         (const double*)v_in, (double*)v_out, nvec, invG);
    else
      hipLaunchKernelGGL((k_shift_sector<I, MM, FM, double2>), grid, blk, 0, st, A, T, other_core->d,
                         other_core->phase, (const I*)other_basis, other_n, other_n_bytes, other_Ns,
                         (const double2*)v_in, (double2*)v_out, nvec, invG);
  });
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_get_amp(qsx_core* core, void* states, void* out, int out_dtype, int64_t n, int mode,
                           void* stream) {
  if (!core || (mode != 0 && mode != 1) || out_dtype < QSX_F32 || out_dtype > QSX_C128) {
    qsx_set_error("qsx_get_amp: bad arguments (mode 0 = representative, 1 = full_basis)");
    return QSX_ERR_ARG;
  }
  if ((out_dtype == QSX_F32 || out_dtype == QSX_F64) && !chars_real(core)) {
    qsx_set_error("real output for a sector with complex characters");
    return QSX_ERR_COMPLEX;
  }
  if (n <= 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)qsx_grid(n, 256)), blk(256);
  const double invG = 1.0 / (double)core->h.G;
  QSX_SEC_DISPATCH(core, {
    switch (out_dtype) {
      case QSX_F32: hipLaunchKernelGGL((k_get_amp<I, MM, FM, float>), grid, blk, 0, st, core->d, core->phase, (I*)states, (float*)out, n, mode, invG); break;
      case QSX_F64: hipLaunchKernelGGL((k_get_amp<I, MM, FM, double>), grid, blk, 0, st, core->d, core->phase, (I*)states, (double*)out, n, mode, invG); break;
      case QSX_C64: hipLaunchKernelGGL((k_get_amp<I, MM, FM, float2>), grid, blk, 0, st, core->d, core->phase, (I*)states, (float2*)out, n, mode, invG); break;
      default: hipLaunchKernelGGL((k_get_amp<I, MM, FM, double2>), grid, blk, 0, st, core->d, core->phase, (I*)states, (double2*)out, n, mode, invG); break;
    }
  });
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_project_from_dense(qsx_core* core, const void* basis, const void* n, int n_bytes, int64_t Ns,
                                      const void* v0, void* v_out, int dtype, int64_t nvec, int pcon,
                                      int64_t Ns_full, void* stream) {
  if (!core || nvec < 1 || dtype < QSX_F32 || dtype > QSX_C128) { qsx_set_error("qsx_project_from_dense: bad arguments"); return QSX_ERR_ARG; }
  int rc = check_pcon(core, pcon, Ns_full);
  if (rc) return rc;
  if ((dtype == QSX_F32 || dtype == QSX_F64) && !chars_real(core)) {
    qsx_set_error("real vectors for a sector with complex characters");
    return QSX_ERR_COMPLEX;
  }
  if (Ns == 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)qsx_grid(Ns, 256)), blk(256);
#define QSX_PF(VT_, PC_)                                                                                        \
  hipLaunchKernelGGL((k_project_from<I, MM, FM, PC_, VT_>), grid, blk, 0, st, core->d, core->phase, core->rank_d, \
                     (const I*)basis, n, n_bytes, Ns, (const VT_*)v0, (VT_*)v_out, nvec)
  QSX_SEC_DISPATCH(core, {
    switch (dtype) {
      case QSX_F32: if (pcon) QSX_PF(float, 1); else QSX_PF(float, 0); break;
      case QSX_F64: if (pcon) QSX_PF(double, 1); else QSX_PF(double, 0); break;
      case QSX_C64: if (pcon) QSX_PF(float2, 1); else QSX_PF(float2, 0); break;
      default: if (pcon) QSX_PF(double2, 1); else QSX_PF(double2, 0); break;
    }
  });
#undef QSX_PF
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_project_to_dense(qsx_core* core, const void* basis, const void* n, int n_bytes, int64_t Ns,
                                    const void* v0, void* v_out, int dtype, int64_t nvec, int pcon, int64_t Ns_full,
                                    void* stream) {
  if (!core || nvec < 1 || dtype < QSX_F32 || dtype > QSX_C128) { qsx_set_error("qsx_project_to_dense: bad arguments"); return QSX_ERR_ARG; }
  int rc = check_pcon(core, pcon, Ns_full);
  if (rc) return rc;
  if ((dtype == QSX_F32 || dtype == QSX_F64) && !chars_real(core)) {
    qsx_set_error("real vectors for a sector with complex characters");
    return QSX_ERR_COMPLEX;
  }
  if (Ns == 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)qsx_grid(Ns, 256)), blk(256);
#define QSX_PT(VT_, PC_)                                                                                      \
  hipLaunchKernelGGL((k_project_to<I, MM, FM, PC_, VT_>), grid, blk, 0, st, core->d, core->phase, core->rank_d, \
                     (const I*)basis, n, n_bytes, Ns, (const VT_*)v0, (VT_*)v_out, nvec)
  QSX_SEC_DISPATCH(core, {
    switch (dtype) {
      case QSX_F32: if (pcon) QSX_PT(float, 1); else QSX_PT(float, 0); break;
      case QSX_F64: if (pcon) QSX_PT(double, 1); else QSX_PT(double, 0); break;
      case QSX_C64: if (pcon) QSX_PT(float2, 1); else QSX_PT(float2, 0); break;
      default: if (pcon) QSX_PT(double2, 1); else QSX_PT(double2, 0); break;
    }
  });
#undef QSX_PT
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// Two calls: with indices == NULL the number of stored entries goes to
// *nnz_out; then indptr (Ns + 1), indices and data (nnz entries) are written.
extern "C" int qsx_get_proj(qsx_core* core, const void* basis, const void* n, int n_bytes, int64_t Ns, int pcon,
                            int64_t Ns_full, int dtype, int idx_bytes, void* indptr, void* indices, void* data,
                            int64_t* nnz_out, void* stream) {
  if (!core || !nnz_out || (idx_bytes != 4 && idx_bytes != 8) || dtype < QSX_F32 || dtype > QSX_C128) {
    qsx_set_error("qsx_get_proj: bad arguments");
    return QSX_ERR_ARG;
  }
  int rc = check_pcon(core, pcon, Ns_full);
  if (rc) return rc;
  if ((dtype == QSX_F32 || dtype == QSX_F64) && !chars_real(core)) {
    qsx_set_error("real projector for a sector with complex characters");
    return QSX_ERR_COMPLEX;
  }
  hipStream_t st = (hipStream_t)stream;
  if (Ns == 0) {
    *nnz_out = 0;
    if (indptr) {
      if (idx_bytes == 4) { int32_t z = 0; QSX_CHECK_HIP(hipMemcpyAsync(indptr, &z, 4, hipMemcpyHostToDevice, st)); }
      else { int64_t z = 0; QSX_CHECK_HIP(hipMemcpyAsync(indptr, &z, 8, hipMemcpyHostToDevice, st)); }
      QSX_CHECK_HIP(hipStreamSynchronize(st));
    }
    return QSX_OK;
  }
  int64_t *cnt = nullptr, *offs = nullptr, *rows = nullptr;
  double2* vals = nullptr;
  void* tmp = nullptr;
  const dim3 grid((unsigned)qsx_grid(Ns, 256)), blk(256);
  auto cleanup = [&]() {
    if (cnt) (void)hipFree(cnt);
    if (offs) (void)hipFree(offs);
    if (rows) (void)hipFree(rows);
    if (vals) (void)hipFree(vals);
    if (tmp) (void)hipFree(tmp);
  };
#define QSX_GP(PASS_)                                                                                          \
  QSX_SEC_DISPATCH(core, {                                                                                     \
    if (pcon) hipLaunchKernelGGL((k_get_proj<I, MM, FM, 1>), grid, blk, 0, st, core->d, core->phase,            \
                                 core->rank_d, (const I*)basis, n, n_bytes, Ns, PASS_, cnt, offs, rows, vals);  \
    else hipLaunchKernelGGL((k_get_proj<I, MM, FM, 0>), grid, blk, 0, st, core->d, core->phase, core->rank_d,  \
                            (const I*)basis, n, n_bytes, Ns, PASS_, cnt, offs, rows, vals);                     \
  })
  hipError_t e = hipMalloc(&cnt, sizeof(int64_t) * Ns);
  if (e == hipSuccess) e = hipMalloc(&offs, sizeof(int64_t) * Ns);
  if (e != hipSuccess) { cleanup(); qsx_set_error("qsx_get_proj: allocation failed"); return QSX_ERR_HIP; }
  QSX_GP(0);
  size_t tb = 0;
  e = hipGetLastError();
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, offs, (int)Ns, st);
  if (e == hipSuccess) e = hipMalloc(&tmp, tb + 16);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, offs, (int)Ns, st);
  int64_t lc = 0, lo = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&lc, cnt + Ns - 1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&lo, offs + Ns - 1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) { cleanup(); qsx_set_error(std::string("qsx_get_proj: ") + hipGetErrorString(e)); return QSX_ERR_HIP; }
  const int64_t nnz = lc + lo;
  *nnz_out = nnz;
  if (!indices) { cleanup(); return QSX_OK; }
  if (!indptr || !data || (idx_bytes == 4 && nnz > INT32_MAX)) {
    cleanup();
    qsx_set_error("qsx_get_proj: null output or 32-bit indices too small");
    return QSX_ERR_ARG;
  }
  e = hipMalloc(&rows, sizeof(int64_t) * (nnz + 1));
  if (e == hipSuccess) e = hipMalloc(&vals, sizeof(double2) * (nnz + 1));
  if (e != hipSuccess) { cleanup(); qsx_set_error("qsx_get_proj: allocation failed"); return QSX_ERR_HIP; }
  QSX_GP(1);
#undef QSX_GP
  hipLaunchKernelGGL(k_proj_write, dim3((unsigned)qsx_grid(Ns + 1, 256)), blk, 0, st, Ns, cnt, offs, rows, vals,
                     indptr, indices, data, idx_bytes, dtype);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) { qsx_set_error(std::string("qsx_get_proj: ") + hipGetErrorString(e)); return QSX_ERR_HIP; }
  return QSX_OK;
}
