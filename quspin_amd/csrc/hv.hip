// Operator lists, fused H.v, per-term inplace_op, COO op and CSR assembly.
//
// Replaces ``_core.inplace_op`` (base_general.py:395-408), ``_core.op``
// (base_general.py:324-336) and the COO->CSR assembly of ``_make_matrix``
// (src/quspin/basis/base.py:265-304).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "hv_impl.h"

namespace {

int op_code(char ch) {
  switch (ch) {
    case 'I': return OP_I;
    case 'z': return OP_Z;
    case 'n': return OP_N;
    case '+': return OP_P;
    case '-': return OP_M;
    case 'x': return OP_X;
    case 'y': return OP_Y;
    default: return -1;
  }
}

bool is_flip(int c) { return c == OP_P || c == OP_M || c == OP_X || c == OP_Y; }

// Compile opstr/indx (character k acts on site indx[k]; rightmost first) into
// a TermDev; hc = 1 builds the Hermitian conjugate (reversed order, + <-> -,
// conj(J)).
int compile_term(const qsx_core* c, const char* opstr, const int32_t* indx, int n_op, double jr, double ji, int hc,
                 TermDev& T, int& n_y) {
  if (n_op > QSX_MAX_OPS) { qsx_set_error("operator string longer than 16 characters"); return QSX_ERR_UNSUPPORTED; }
  std::memset(&T, 0, sizeof(T));
  const int N = c->h.N;
  T.nops = n_op;
  T.jr = jr;
  T.ji = hc ? -ji : ji;
  n_y = 0;
  for (int k = 0; k < n_op; ++k) {
    // application order: original -> characters n_op-1 ... 0; hc -> 0 ... n_op-1
    const int src = hc ? k : (n_op - 1 - k);
    int code = op_code(opstr[src]);
    if (code < 0) {
      qsx_set_error(std::string("unrecognized characters {'") + opstr[src] + "'} in operator string.");
      return QSX_ERR_ARG;
    }
    if (hc) { if (code == OP_P) code = OP_M; else if (code == OP_M) code = OP_P; }
    const int site = indx[src];
    if (site < 0 || site >= N) { qsx_set_error("values in indx falls outside of system"); return QSX_ERR_ARG; }
    T.code[k] = (uint8_t)code;
    T.pos[k] = (uint8_t)(N - 1 - site);
    if (is_flip(code)) T.flip ^= 1ull << (N - 1 - site);
    if (code == OP_Y) ++n_y;
  }
  if (T.flip) {
    T.lo = (int16_t)__builtin_ctzll(T.flip);
    T.hi = (int16_t)(63 - __builtin_clzll(T.flip));
  }
  // bit-parallel form (spin / hard-core bosons, every site at most once)
  if (!c->h.fermion) {
    uint64_t seen = 0;
    bool distinct = true;
    for (int k = 0; k < n_op; ++k) {
      const uint64_t b = 1ull << T.pos[k];
      if (seen & b) distinct = false;
      seen |= b;
    }
    if (distinct) {
      T.fast = 1;
      T.mag0 = 1.0;
      for (int k = 0; k < n_op; ++k) {
        const uint64_t b = 1ull << T.pos[k];
        switch (T.code[k]) {
          case OP_Z: T.zmask |= b; T.nz += 1; T.mag0 *= 0.5; break;
          case OP_N: T.req_mask |= b; T.req_val |= b; break;
          case OP_P: T.req_mask |= b; break;
          case OP_M: T.req_mask |= b; T.req_val |= b; break;
          case OP_X: T.mag0 *= 0.5; break;
          case OP_Y: T.ymask |= b; T.ny += 1; T.mag0 *= 0.5; break;
          default: break;
        }
      }
      double re = T.jr * T.mag0, im = T.ji * T.mag0;
      if (T.nz & 1) { re = -re; im = -im; }
      switch ((3 * T.ny) & 3) {  // times i^(3 ny)
        case 1: { const double t = re; re = -im; im = t; } break;
        case 2: re = -re; im = -im; break;
        case 3: { const double t = re; re = im; im = -t; } break;
        default: break;
      }
      T.cr = re;
      T.ci = im;
      T.smask = T.zmask ^ T.ymask;
    }
  }
  return QSX_OK;
}

bool mergeable(const TermDev& a, const TermDev& b) {
  return a.fast && b.fast && !a.has2 && !b.has2 && a.flip == b.flip && a.req_mask == b.req_mask &&
         a.req_val != b.req_val && a.zmask == b.zmask && a.ymask == b.ymask && a.nz == b.nz && a.ny == b.ny &&
         a.mag0 == b.mag0 && a.jr == b.jr && a.ji == b.ji;
}

// Merge off-diagonal terms that differ only in the bits they require (same
// flip, amplitude form and coupling): at most one of the two applies to a
// state and both reach s ^ flip with the same element, so the H.v checks
// both patterns and resolves the target once.  The pairing is applied to the
// term list and its Hermitian-conjugate list together (they stay aligned).
void merge_pairs(std::vector<TermDev>& off, std::vector<TermDev>& hc) {
  const size_t n = off.size();
  std::vector<char> gone(n, 0);
  for (size_t i = 0; i < n; ++i) {
    if (gone[i] || !off[i].fast) continue;
    for (size_t j = i + 1; j < n; ++j) {
      if (gone[j] || !mergeable(off[i], off[j]) || !mergeable(hc[i], hc[j])) continue;
      off[i].req_val2 = off[j].req_val;
      off[i].has2 = 1;
      hc[i].req_val2 = hc[j].req_val;
      hc[i].has2 = 1;
      gone[j] = 1;
      break;
    }
  }
  size_t k = 0;
  for (size_t i = 0; i < n; ++i) {
    if (gone[i]) continue;
    off[k] = off[i];
    hc[k] = hc[i];
    ++k;
  }
  off.resize(k);
  hc.resize(k);
}

}  // namespace

int qsx_compile_term(const qsx_core* c, const char* opstr, const int32_t* indx, int n_op, double jr, double ji, int hc,
                     TermDev& T, int& n_y) {
  return compile_term(c, opstr, indx, n_op, jr, ji, hc, T, n_y);
}

struct qsx_oplist_full {
  qsx_oplist base;
  TermDev* off_hc = nullptr;  // device, Hermitian-conjugate off-diagonal terms
  std::vector<TermDev> off_hc_h;
  int real_ok = 1;            // all factors real for real dtypes
  int symmetric = 1;          // list commutes with the symmetry group (gather form exact)
  int all_fast = 0;           // every off-diagonal term (and its conjugate) in the bit-parallel form
};

extern "C" int qsx_oplist_create(qsx_core* c, int n_terms, const char* opstrs, const int32_t* lens,
                                 const int32_t* indx, const double* J, int symmetric, qsx_oplist** out) {
  if (!c || !out || n_terms < 0 || n_terms > QSX_MAX_TERMS) {
    qsx_set_error("qsx_oplist_create: bad arguments (at most 4096 terms)");
    return QSX_ERR_ARG;
  }
  auto* L = new qsx_oplist_full();
  L->symmetric = symmetric ? 1 : 0;
  int pos_s = 0, pos_i = 0;
  for (int t = 0; t < n_terms; ++t) {
    TermDev T, Th;
    int ny = 0, ny2 = 0;
    int rc = compile_term(c, opstrs + pos_s, indx + pos_i, lens[t], J[2 * t], J[2 * t + 1], 0, T, ny);
    if (rc == QSX_OK) rc = compile_term(c, opstrs + pos_s, indx + pos_i, lens[t], J[2 * t], J[2 * t + 1], 1, Th, ny2);
    if (rc != QSX_OK) { delete L; return rc; }
    pos_s += lens[t];
    pos_i += lens[t];
    if (J[2 * t + 1] != 0.0 || (ny & 1)) L->real_ok = 0;
    if (ny > 0 || J[2 * t + 1] != 0.0) L->base.complex_me = 1;
    if (T.flip == 0) L->base.diag_h.push_back(T);
    else { L->base.off_h.push_back(T); L->off_hc_h.push_back(Th); }
  }
  merge_pairs(L->base.off_h, L->off_hc_h);
  L->all_fast = 1;
  for (size_t i = 0; i < L->base.off_h.size(); ++i)
    if (!L->base.off_h[i].fast || !L->off_hc_h[i].fast) L->all_fast = 0;
  L->base.n_off = (int)L->base.off_h.size();
  L->base.n_diag = (int)L->base.diag_h.size();
  auto up = [](std::vector<TermDev>& v, TermDev** d) -> hipError_t {
    if (v.empty()) return hipSuccess;
    hipError_t e = hipMalloc(d, sizeof(TermDev) * v.size());
    if (e == hipSuccess) e = hipMemcpy(*d, v.data(), sizeof(TermDev) * v.size(), hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = up(L->base.off_h, &L->base.off);
  if (e == hipSuccess) e = up(L->off_hc_h, &L->off_hc);
  if (e == hipSuccess) e = up(L->base.diag_h, &L->base.diag);
  if (e != hipSuccess) {
    qsx_set_error(std::string("qsx_oplist_create: ") + hipGetErrorString(e));
    delete L;
    return QSX_ERR_HIP;
  }
  *out = &L->base;
  return QSX_OK;
}

extern "C" void qsx_oplist_destroy(qsx_oplist* l) {
  if (!l) return;
  auto* L = reinterpret_cast<qsx_oplist_full*>(l);
  if (L->base.off) (void)hipFree(L->base.off);
  if (L->off_hc) (void)hipFree(L->off_hc);
  if (L->base.diag) (void)hipFree(L->base.diag);
  delete L;
}

extern "C" int qsx_oplist_info(const qsx_oplist* l, int* n_off, int* n_diag, int* complex_me) {
  if (!l) return QSX_ERR_ARG;
  *n_off = l->n_off;
  *n_diag = l->n_diag;
  *complex_me = l->complex_me;
  return QSX_OK;
}

static bool core_chars_real(const qsx_core* c) {
  for (const auto& p : c->phase_h) if (p.y != 0.0) return false;
  return true;
}

static int pick_lookup(const qsx_core* c, int64_t Ns) {
  if (c->h.nmaps == 0 && c->rank_ok && c->rank_h.Ns == Ns) return LOOKUP_RANK;
  return LOOKUP_SEARCH;
}

static int dispatch_hv(const HvArgs& A, int dtype, int N, int nmaps, bool fm, int lookup, int scatter,
                       hipStream_t st) {
  switch (dtype) {
    case QSX_F32: return qsx_hv_launch_f32(A, N, nmaps, fm, lookup, scatter, st);
    case QSX_F64: return qsx_hv_launch_f64(A, N, nmaps, fm, lookup, scatter, st);
    case QSX_C64: return qsx_hv_launch_c64(A, N, nmaps, fm, lookup, scatter, st);
    case QSX_C128: return qsx_hv_launch_c128(A, N, nmaps, fm, lookup, scatter, st);
    default: qsx_set_error("unknown dtype code"); return QSX_ERR_ARG;
  }
}

// Launch v_out = beta v_out + (a_diag D + shift) v_in + a_off Op v_in.  A
// carries everything except the off-diagonal terms.  Symmetric lists (and
// every transposed product) use the one-launch gather; otherwise the diagonal
// launch is followed by the scatter launch with the original terms.
static int hv_launch(const qsx_core* c, const qsx_oplist_full* L, HvArgs A, int transposed, int conjugated,
                     int dtype, hipStream_t st) {
  const int lookup = A.table ? LOOKUP_TABLE : pick_lookup(c, A.Ns);
  const int N = c->h.N, nm = c->h.nmaps;
  const bool fm = c->h.fermion != 0;
  const bool scatter = nm > 0 && !L->symmetric && !transposed && L->base.n_off > 0;
  // binomial rows 0..N, rounded up to an even count (16-byte aligned LDS after them)
  A.n_sb = ((std::min(65, c->rank_h.N + 1) * 33) + 1) & ~1;
  if (!scatter) {
    A.off = transposed ? L->base.off : L->off_hc;
    A.n_off = L->base.n_off;
    A.off_conj = (transposed == conjugated) ? 1 : 0;
    A.fast_rows = (L->all_fast && A.n_diag == 0 && !fm && c->rank_h.mode == 0 &&
                   ((lookup == LOOKUP_TABLE && A.tlut) || lookup == LOOKUP_RANK)) ? 1 : 0;
    return dispatch_hv(A, dtype, N, nm, fm, lookup, 0, st);
  }
  A.fast_rows = 0;
  HvArgs D = A;
  D.off = nullptr;
  D.n_off = 0;
  int rc = dispatch_hv(D, dtype, N, nm, fm, lookup, 0, st);
  if (rc) return rc;
  HvArgs S = A;
  S.off = L->base.off;
  S.n_off = L->base.n_off;
  S.off_conj = conjugated ? 1 : 0;
  S.dg = nullptr;
  S.n_diag = 0;
  S.diag_arr = nullptr;
  return dispatch_hv(S, dtype, N, nm, fm, lookup, 1, st);
}

// v_out = beta*v_out + a_diag*(D + diag_arr) v_in + a_off*Op v_in
// (Op = sum of off-diagonal terms, transposed/conjugated per flags; D = sum of
// diagonal terms, conjugated iff diag_conj).
extern "C" int qsx_hv(qsx_core* c, qsx_oplist* l, const void* basis, const void* nrm, int n_bytes, int64_t Ns,
                      const int64_t* begin, const int64_t* end, int N_p, const void* v_in, void* v_out, int dtype,
                      int64_t nvec, int transposed, int conjugated, int diag_conj, double a_off_re,
                      double a_off_im, double a_diag_re, double a_diag_im, int beta, const void* diag_arr,
                      int diag_arr_cplx, void* stream) {
  if (!c || !l || nvec < 1 || Ns < 0) { qsx_set_error("qsx_hv: bad arguments"); return QSX_ERR_ARG; }
  if (Ns == 0) return QSX_OK;
  auto* L = reinterpret_cast<qsx_oplist_full*>(l);
  const bool real = (dtype == QSX_F32 || dtype == QSX_F64);
  if (real && (!L->real_ok || !core_chars_real(c) || a_off_im != 0.0 || a_diag_im != 0.0 ||
               (diag_arr && diag_arr_cplx))) {
    qsx_set_error("real dtype requested for an operator with complex matrix elements");
    return QSX_ERR_COMPLEX;
  }
  if (c->h.nmaps > 0 && n_bytes != 1 && n_bytes != 2 && n_bytes != 4) {
    qsx_set_error("qsx_hv: n_bytes must be 1, 2 or 4");
    return QSX_ERR_ARG;
  }
  HvArgs A;
  std::memset(&A, 0, sizeof(A));
  A.C = c->d;
  A.ph = c->phase;
  A.R = c->rank_d;
  A.dg = L->base.diag;
  A.n_diag = L->base.n_diag;
  A.diag_conj = diag_conj ? 1 : 0;
  A.diag_arr = diag_arr;
  A.diag_arr_cplx = diag_arr_cplx;
  A.basis = basis;
  A.nrm = nrm;
  A.n_bytes = n_bytes;
  A.begin = begin;
  A.end = end;
  A.N_p = N_p;
  A.v_in = v_in;
  A.v_out = v_out;
  A.Ns = Ns;
  A.nvec = nvec;
  A.alpha_off = make_double2(a_off_re, a_off_im);
  A.alpha_diag = make_double2(a_diag_re, a_diag_im);
  A.beta = beta ? 1 : 0;
  return hv_launch(c, L, A, transposed, conjugated, dtype, (hipStream_t)stream);
}

// Reference-compatible single term: v_out += J Op v_in (Op^T / Op^* per flags).
extern "C" int qsx_inplace_op(qsx_core* c, const void* v_in, void* v_out, int dtype, int64_t nvec, int conjugated,
                              int transposed, const char* opstr, const int32_t* indx, int n_op, double J_re,
                              double J_im, const void* basis, const void* nrm, int n_bytes, int64_t Ns,
                              const int64_t* begin, const int64_t* end, int N_p, void* stream) {
  if (!c) { qsx_set_error("qsx_inplace_op: null core"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  TermDev T, Th;
  int ny = 0, ny2 = 0;
  int rc = compile_term(c, opstr, indx, n_op, J_re, J_im, 0, T, ny);
  if (rc == QSX_OK) rc = compile_term(c, opstr, indx, n_op, J_re, J_im, 1, Th, ny2);
  if (rc != QSX_OK) return rc;
  if (Ns == 0) return QSX_OK;
  const bool real = (dtype == QSX_F32 || dtype == QSX_F64);
  if (real && (J_im != 0.0 || (ny & 1) || !core_chars_real(c))) {
    qsx_set_error("real dtype requested for an operator with complex matrix elements");
    return QSX_ERR_COMPLEX;
  }
  // a single term is handled as a one-term list that is not assumed to
  // commute with the group: transposed -> gather with the term, otherwise
  // scatter (the reference's v_out[row] += ME v_in[col])
  qsx_oplist_full one;
  one.symmetric = 0;
  TermDev* d = nullptr;
  QSX_CHECK_HIP(hipMallocAsync((void**)&d, 2 * sizeof(TermDev), st));
  QSX_CHECK_HIP(hipMemcpyAsync(d, &T, sizeof(TermDev), hipMemcpyHostToDevice, st));
  QSX_CHECK_HIP(hipMemcpyAsync(d + 1, &Th, sizeof(TermDev), hipMemcpyHostToDevice, st));
  if (T.flip == 0) {
    one.base.diag = d; one.base.n_diag = 1;
  } else {
    one.base.off = d; one.off_hc = d + 1; one.base.n_off = 1;
  }
  HvArgs A;
  std::memset(&A, 0, sizeof(A));
  A.C = c->d;
  A.ph = c->phase;
  A.R = c->rank_d;
  A.dg = one.base.diag; A.n_diag = one.base.n_diag; A.diag_conj = conjugated ? 1 : 0;
  A.basis = basis; A.nrm = nrm; A.n_bytes = n_bytes; A.begin = begin; A.end = end; A.N_p = N_p;
  A.v_in = v_in; A.v_out = v_out; A.Ns = Ns; A.nvec = nvec;
  A.alpha_off = make_double2(1.0, 0.0);
  A.alpha_diag = make_double2(1.0, 0.0);
  A.beta = 1;
  rc = hv_launch(c, &one, A, transposed, conjugated, dtype, st);
  one.base.off = nullptr; one.off_hc = nullptr; one.base.diag = nullptr;
  (void)hipFreeAsync(d, st);
  return rc;
}

// ---------------------------------------------------------------------------
// COO of one term, one element per basis state (column): ``_core.op``.
// ---------------------------------------------------------------------------
template <typename I, int M, bool FERM, int LK>
__global__ void __launch_bounds__(256) k_op_coo(HvArgs A, const TermDev* __restrict__ T, void* row, void* col,
                                                int idx_bytes, double2* __restrict__ me) {
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.Ns) return;
  const I s = ((const I*)A.basis)[i];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, i));
  int64_t rr = i;
  double2 v = make_double2(0.0, 0.0);
  const RowRank rk = row_rank_if<I, LK>(A, sb, s);
  term_element<I, M, FERM, LK, true>(A, sb, rk, T, s, rsq, 0, [&](int64_t c, double2 coef) { rr = c; v = coef; });
  if (idx_bytes == 4) { ((int32_t*)row)[i] = (int32_t)rr; ((int32_t*)col)[i] = (int32_t)i; }
  else { ((int64_t*)row)[i] = rr; ((int64_t*)col)[i] = i; }
  me[i] = v;
}

template <typename I>
static void launch_op_coo(const HvArgs& A, const TermDev* T, void* row, void* col, int idx_bytes, double2* me,
                          int nmaps, bool fm, int lookup, hipStream_t st) {
  const dim3 grid((unsigned)qsx_grid(A.Ns, 256)), blk(256);
  if (nmaps == 0) {
    if (lookup == LOOKUP_RANK) hipLaunchKernelGGL((k_op_coo<I, 0, false, LOOKUP_RANK>), grid, blk, 0, st, A, T, row, col, idx_bytes, me);
    else hipLaunchKernelGGL((k_op_coo<I, 0, false, LOOKUP_SEARCH>), grid, blk, 0, st, A, T, row, col, idx_bytes, me);
  } else if (fm) {
    QSX_HV_MSWITCH(nmaps, hipLaunchKernelGGL((k_op_coo<I, MM, true, LOOKUP_SEARCH>), grid, blk, 0, st, A, T, row, col, idx_bytes, me));
  } else {
    QSX_HV_MSWITCH(nmaps, hipLaunchKernelGGL((k_op_coo<I, MM, false, LOOKUP_SEARCH>), grid, blk, 0, st, A, T, row, col, idx_bytes, me));
  }
}

// me_out: complex128 (double2) per basis state; zero where the element
// vanishes.  The caller casts and masks (base_general.py:338-350).
extern "C" int qsx_op(qsx_core* c, void* row, void* col, int idx_bytes, void* me_out, const char* opstr,
                      const int32_t* indx, int n_op, double J_re, double J_im, const void* basis, const void* nrm,
                      int n_bytes, int64_t Ns, const int64_t* begin, const int64_t* end, int N_p, void* stream) {
  if (!c || (idx_bytes != 4 && idx_bytes != 8)) { qsx_set_error("qsx_op: bad arguments"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  TermDev T;
  int ny = 0;
  int rc = compile_term(c, opstr, indx, n_op, J_re, J_im, 0, T, ny);
  if (rc != QSX_OK) return rc;
  if (Ns == 0) return QSX_OK;
  TermDev* d = nullptr;
  QSX_CHECK_HIP(hipMallocAsync((void**)&d, sizeof(TermDev), st));
  QSX_CHECK_HIP(hipMemcpyAsync(d, &T, sizeof(TermDev), hipMemcpyHostToDevice, st));
  HvArgs A;
  std::memset(&A, 0, sizeof(A));
  A.C = c->d; A.ph = c->phase; A.R = c->rank_d;
  A.basis = basis; A.nrm = nrm; A.n_bytes = n_bytes; A.begin = begin; A.end = end; A.N_p = N_p; A.Ns = Ns;
  const int lk = pick_lookup(c, Ns);
  if (c->h.N <= 32) launch_op_coo<uint32_t>(A, d, row, col, idx_bytes, (double2*)me_out, c->h.nmaps, c->h.fermion != 0, lk, st);
  else launch_op_coo<uint64_t>(A, d, row, col, idx_bytes, (double2*)me_out, c->h.nmaps, c->h.fermion != 0, lk, st);
  QSX_LAUNCH_CHECK();
  (void)hipFreeAsync(d, st);
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// CSR assembly: rows gathered with the Hermitian-conjugate terms, columns
// sorted, duplicates summed, zeros dropped (scipy canonical CSR).
// ---------------------------------------------------------------------------
template <typename I, int M, bool FERM, int LK, int PASS>
__global__ void __launch_bounds__(256) k_csr_rows(HvArgs A, int64_t* __restrict__ cnt, const int64_t* __restrict__ offs,
                                                  int64_t* __restrict__ rc, double2* __restrict__ rv) {
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.Ns) return;
  const I s = ((const I*)A.basis)[r];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
  const double2 d = row_diag<I, M, FERM, LK, true>(A, r, s);
  const RowRank rr = row_rank_if<I, LK>(A, sb, s);
  int64_t k = 0;
  const int64_t o = (PASS == 1) ? offs[r] : 0;
  if (d.x != 0.0 || d.y != 0.0) {
    if constexpr (PASS == 1) { rc[o + k] = r; rv[o + k] = d; }
    ++k;
  }
  for (int t = 0; t < A.n_off; ++t) {
    term_element<I, M, FERM, LK, true>(A, sb, rr, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
      coef = cmul(coef, A.alpha_off);
      if constexpr (PASS == 1) { rc[o + k] = c; rv[o + k] = coef; }
      ++k;
    });
  }
  if constexpr (PASS == 0) cnt[r] = k;
}

// per row: insertion sort by column, merge duplicates, drop zeros
__global__ void k_csr_canon(int64_t Ns, const int64_t* __restrict__ offs, const int64_t* __restrict__ cnt,
                            int64_t* __restrict__ rc, double2* __restrict__ rv, int64_t* __restrict__ fcnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Ns) return;
  const int64_t o = offs[r], n = cnt[r];
  for (int64_t i = 1; i < n; ++i) {
    const int64_t kc = rc[o + i];
    const double2 kv = rv[o + i];
    int64_t j = i - 1;
    while (j >= 0 && rc[o + j] > kc) { rc[o + j + 1] = rc[o + j]; rv[o + j + 1] = rv[o + j]; --j; }
    rc[o + j + 1] = kc;
    rv[o + j + 1] = kv;
  }
  int64_t w = 0;
  for (int64_t i = 0; i < n;) {
    const int64_t cc = rc[o + i];
    double2 acc = rv[o + i];
    int64_t j = i + 1;
    while (j < n && rc[o + j] == cc) { acc = cadd(acc, rv[o + j]); ++j; }
    if (acc.x != 0.0 || acc.y != 0.0) { rc[o + w] = cc; rv[o + w] = acc; ++w; }
    i = j;
  }
  fcnt[r] = w;
}

__global__ void k_csr_compact(int64_t Ns, const int64_t* __restrict__ offs, const int64_t* __restrict__ fcnt,
                              const int64_t* __restrict__ foffs, const int64_t* __restrict__ rc,
                              const double2* __restrict__ rv, void* indptr, void* indices, void* data,
                              int idx_bytes, int dtype) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > Ns) return;
  if (r == Ns) {
    const int64_t tot = (Ns > 0) ? foffs[Ns - 1] + fcnt[Ns - 1] : 0;
    if (idx_bytes == 4) ((int32_t*)indptr)[Ns] = (int32_t)tot; else ((int64_t*)indptr)[Ns] = tot;
    return;
  }
  const int64_t o = offs[r], fo = foffs[r], n = fcnt[r];
  if (idx_bytes == 4) ((int32_t*)indptr)[r] = (int32_t)fo; else ((int64_t*)indptr)[r] = fo;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t cc = rc[o + i];
    const double2 v = rv[o + i];
    if (idx_bytes == 4) ((int32_t*)indices)[fo + i] = (int32_t)cc; else ((int64_t*)indices)[fo + i] = cc;
    switch (dtype) {
      case QSX_F32: ((float*)data)[fo + i] = (float)v.x; break;
      case QSX_F64: ((double*)data)[fo + i] = v.x; break;
      case QSX_C64: ((float2*)data)[fo + i] = make_float2((float)v.x, (float)v.y); break;
      default: ((double2*)data)[fo + i] = v; break;
    }
  }
}

struct qsx_csr_builder {
  int64_t Ns = 0, nnz = 0;
  int64_t *offs = nullptr, *cnt = nullptr, *fcnt = nullptr, *foffs = nullptr, *rc = nullptr;
  double2* rv = nullptr;
  int complex_vals = 0;
};

template <typename I, int M, bool FERM, int LK>
static void launch_csr_rows(const HvArgs& A, int pass, qsx_csr_builder* B, hipStream_t st) {
  const dim3 grid((unsigned)qsx_grid(A.Ns, 256)), blk(256);
  if (pass == 0) hipLaunchKernelGGL((k_csr_rows<I, M, FERM, LK, 0>), grid, blk, 0, st, A, B->cnt, B->offs, B->rc, B->rv);
  else hipLaunchKernelGGL((k_csr_rows<I, M, FERM, LK, 1>), grid, blk, 0, st, A, B->cnt, B->offs, B->rc, B->rv);
}

template <typename I>
static void csr_rows_dispatch(const HvArgs& A, int pass, qsx_csr_builder* B, int nmaps, bool fm, int lk, hipStream_t st) {
  if (nmaps == 0) {
    if (lk == LOOKUP_RANK) launch_csr_rows<I, 0, false, LOOKUP_RANK>(A, pass, B, st);
    else launch_csr_rows<I, 0, false, LOOKUP_SEARCH>(A, pass, B, st);
  } else if (fm) {
    QSX_HV_MSWITCH(nmaps, (launch_csr_rows<I, MM, true, LOOKUP_SEARCH>(A, pass, B, st)));
  } else {
    QSX_HV_MSWITCH(nmaps, (launch_csr_rows<I, MM, false, LOOKUP_SEARCH>(A, pass, B, st)));
  }
}

extern "C" void qsx_csr_builder_destroy(qsx_csr_builder* B) {
  if (!B) return;
  (void)hipFree(B->offs); (void)hipFree(B->cnt); (void)hipFree(B->fcnt); (void)hipFree(B->foffs);
  (void)hipFree(B->rc); (void)hipFree(B->rv);
  delete B;
}

// ---------------------------------------------------------------------------
// Column 1-norms and diagonal of a matrix-free operator without assembling
// it: lane r evaluates row r of the operator's transpose (= column r of the
// operator) with the CSR builder's element code, sums the elements that land
// on one row before taking |.| (as the canonical CSR of the reference holds
// them, summed duplicates), and adds the shift on the diagonal.
// ---------------------------------------------------------------------------
constexpr int NORM_ROW_MAX = 64;  // distinct off-diagonal rows per column

template <typename I, int M, bool FERM, int LK>
__global__ void __launch_bounds__(256) k_norm_cols(HvArgs A, double2 shift, double* __restrict__ colabs,
                                                   double2* __restrict__ dg, int* __restrict__ overflow) {
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.Ns) return;
  const I s = ((const I*)A.basis)[r];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
  double2 d = row_diag<I, M, FERM, LK, true>(A, r, s);
  const RowRank rr = row_rank_if<I, LK>(A, sb, s);
  int64_t cols[NORM_ROW_MAX];
  double2 vals[NORM_ROW_MAX];
  int n = 0;
  bool over = false;
  for (int t = 0; t < A.n_off; ++t) {
    term_element<I, M, FERM, LK, true>(A, sb, rr, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
      if (c == r) { d = cadd(d, coef); return; }
      int k = 0;
      while (k < n && cols[k] != c) ++k;
      if (k < n) vals[k] = cadd(vals[k], coef);
      else if (n < NORM_ROW_MAX) { cols[n] = c; vals[n] = coef; ++n; }
      else over = true;
    });
  }
  dg[r] = d;
  double sum = hypot(d.x + shift.x, d.y + shift.y);
  for (int k = 0; k < n; ++k) sum += hypot(vals[k].x, vals[k].y);
  colabs[r] = sum;
  if (over) *overflow = 1;
}

template <typename I, int M, bool FERM, int LK>
static void launch_norm_cols(const HvArgs& A, double2 shift, double* colabs, double2* dg, int* ovf, hipStream_t st) {
  hipLaunchKernelGGL((k_norm_cols<I, M, FERM, LK>), dim3((unsigned)qsx_grid(A.Ns, 256)), dim3(256), 0, st, A, shift,
                     colabs, dg, ovf);
}

template <typename I>
static void norm_cols_dispatch(const HvArgs& A, double2 shift, double* colabs, double2* dg, int* ovf, int nmaps,
                               bool fm, int lk, hipStream_t st) {
  if (nmaps == 0) {
    if (lk == LOOKUP_RANK) launch_norm_cols<I, 0, false, LOOKUP_RANK>(A, shift, colabs, dg, ovf, st);
    else launch_norm_cols<I, 0, false, LOOKUP_SEARCH>(A, shift, colabs, dg, ovf, st);
  } else if (fm) {
    QSX_HV_MSWITCH(nmaps, (launch_norm_cols<I, MM, true, LOOKUP_SEARCH>(A, shift, colabs, dg, ovf, st)));
  } else {
    QSX_HV_MSWITCH(nmaps, (launch_norm_cols<I, MM, false, LOOKUP_SEARCH>(A, shift, colabs, dg, ovf, st)));
  }
}

// Assemble a*Op (all terms of l; transposed/conjugated per flags) as CSR.
// Returns nnz through *nnz; finish with qsx_csr_builder_finish.
extern "C" int qsx_csr_builder_create(qsx_core* c, qsx_oplist* l, const void* basis, const void* nrm, int n_bytes,
                                      int64_t Ns, const int64_t* begin, const int64_t* end, int N_p, int transposed,
                                      int conjugated, double a_re, double a_im, qsx_csr_builder** out,
                                      int64_t* nnz, void* stream) {
  if (!c || !l || !out || !nnz) { qsx_set_error("qsx_csr_builder_create: bad arguments"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  auto* L = reinterpret_cast<qsx_oplist_full*>(l);
  if (c->h.nmaps > 0 && !L->symmetric && !transposed) {
    qsx_set_error("CSR rows of a list that does not commute with the symmetry group: assemble the transpose");
    return QSX_ERR_UNSUPPORTED;
  }
  auto* B = new qsx_csr_builder();
  B->Ns = Ns;
  B->complex_vals = (!L->real_ok || !core_chars_real(c) || a_im != 0.0) ? 1 : 0;
  *out = B;
  *nnz = 0;
  if (Ns == 0) return QSX_OK;
  HvArgs A;
  std::memset(&A, 0, sizeof(A));
  A.C = c->d; A.ph = c->phase; A.R = c->rank_d;
  A.off = transposed ? L->base.off : L->off_hc;
  A.n_off = L->base.n_off;
  A.off_conj = (transposed == conjugated) ? 1 : 0;
  A.dg = L->base.diag; A.n_diag = L->base.n_diag; A.diag_conj = conjugated ? 1 : 0;
  A.basis = basis; A.nrm = nrm; A.n_bytes = n_bytes; A.begin = begin; A.end = end; A.N_p = N_p; A.Ns = Ns;
  A.alpha_off = make_double2(a_re, a_im);
  A.alpha_diag = make_double2(a_re, a_im);
  const int lk = pick_lookup(c, Ns);
  QSX_CHECK_HIP(hipMalloc(&B->cnt, sizeof(int64_t) * Ns));
  QSX_CHECK_HIP(hipMalloc(&B->offs, sizeof(int64_t) * Ns));
  QSX_CHECK_HIP(hipMalloc(&B->fcnt, sizeof(int64_t) * Ns));
  QSX_CHECK_HIP(hipMalloc(&B->foffs, sizeof(int64_t) * Ns));
  if (c->h.N <= 32) csr_rows_dispatch<uint32_t>(A, 0, B, c->h.nmaps, c->h.fermion != 0, lk, st);
  else csr_rows_dispatch<uint64_t>(A, 0, B, c->h.nmaps, c->h.fermion != 0, lk, st);
  QSX_LAUNCH_CHECK();
  void* tmp = nullptr;
  size_t tb = 0;
  QSX_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, B->cnt, B->offs, (int)Ns, st));
  QSX_CHECK_HIP(hipMalloc(&tmp, tb + 16));
  QSX_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, B->cnt, B->offs, (int)Ns, st));
  int64_t lc = 0, lo = 0;
  QSX_CHECK_HIP(hipMemcpyAsync(&lc, B->cnt + Ns - 1, 8, hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipMemcpyAsync(&lo, B->offs + Ns - 1, 8, hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipStreamSynchronize(st));
  const int64_t raw = lc + lo;
  QSX_CHECK_HIP(hipMalloc(&B->rc, sizeof(int64_t) * (raw + 1)));
  QSX_CHECK_HIP(hipMalloc(&B->rv, sizeof(double2) * (raw + 1)));
  if (c->h.N <= 32) csr_rows_dispatch<uint32_t>(A, 1, B, c->h.nmaps, c->h.fermion != 0, lk, st);
  else csr_rows_dispatch<uint64_t>(A, 1, B, c->h.nmaps, c->h.fermion != 0, lk, st);
  QSX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_csr_canon, dim3((unsigned)qsx_grid(Ns, 256)), dim3(256), 0, st, Ns, B->offs, B->cnt, B->rc,
                     B->rv, B->fcnt);
  QSX_LAUNCH_CHECK();
  QSX_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, B->fcnt, B->foffs, (int)Ns, st));
  QSX_CHECK_HIP(hipMemcpyAsync(&lc, B->fcnt + Ns - 1, 8, hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipMemcpyAsync(&lo, B->foffs + Ns - 1, 8, hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipStreamSynchronize(st));
  (void)hipFree(tmp);
  B->nnz = lc + lo;
  *nnz = B->nnz;
  return QSX_OK;
}

extern "C" int qsx_csr_builder_complex(const qsx_csr_builder* B) { return B ? B->complex_vals : -1; }

extern "C" int qsx_csr_builder_finish(qsx_csr_builder* B, int dtype, int idx_bytes, void* indptr, void* indices,
                                      void* data, void* stream) {
  if (!B || (idx_bytes != 4 && idx_bytes != 8)) { qsx_set_error("qsx_csr_builder_finish: bad arguments"); return QSX_ERR_ARG; }
  hipStream_t st = (hipStream_t)stream;
  if ((dtype == QSX_F32 || dtype == QSX_F64) && B->complex_vals) {
    qsx_set_error("real dtype requested for an operator with complex matrix elements");
    return QSX_ERR_COMPLEX;
  }
  if (B->Ns == 0) {
    if (idx_bytes == 4) { int32_t z = 0; QSX_CHECK_HIP(hipMemcpyAsync(indptr, &z, 4, hipMemcpyHostToDevice, st)); }
    else { int64_t z = 0; QSX_CHECK_HIP(hipMemcpyAsync(indptr, &z, 8, hipMemcpyHostToDevice, st)); }
    QSX_CHECK_HIP(hipStreamSynchronize(st));
    return QSX_OK;
  }
  hipLaunchKernelGGL(k_csr_compact, dim3((unsigned)qsx_grid(B->Ns + 1, 256)), dim3(256), 0, st, B->Ns, B->offs,
                     B->fcnt, B->foffs, B->rc, B->rv, indptr, indices, data, idx_bytes, dtype);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// reusable H.v context (fixed operator, basis, dtype, nvec and flags)
// ---------------------------------------------------------------------------
struct qsx_hv_ctx {
  HvArgs A;  // everything but the off-diagonal terms
  const qsx_core* c;
  const qsx_oplist_full* L;
  int dtype, transposed, conjugated;
  uint64_t owner;  // id of the captures that hold A (renewed when A changes)
  // split-chain plan of schedules 7 and 11 (device allocations owned here;
  // split_z: the plan holds the row-pass buffer of schedule 11)
  std::vector<void*> split_bufs;
  SplitDev* split_d = nullptr;
  bool split_z = false;
  int split_nxt = 0;  // up + crossing terms of the plan
};

static void split_free(qsx_hv_ctx* X) {
  for (void* p : X->split_bufs) (void)hipFree(p);
  X->split_bufs.clear();
  X->split_d = nullptr;
  X->split_z = false;
  X->A.split = nullptr;
  X->A.split_lds = 0;
  X->A.split_nblk = 0;
  X->A.split_nunits = 0;
  X->A.split_nrp = 0;
  X->A.split_lds2 = 0;
  X->A.split_pairs = 0;
}

// Build the split-chain plan of a context (hard-core-boson sector of N <= 32
// bits, no symmetry maps, every off-diagonal term in the bit-parallel form);
// with_z also allocates the row-pass buffer of the two-pass form.
static int split_build(qsx_hv_ctx* X, bool with_z) {
  const qsx_core* c = X->c;
  const qsx_oplist_full* L = X->L;
  const int N = c->h.N;
  if (c->h.nmaps != 0 || c->h.kind != KIND_HCB || N > 32 || !c->rank_ok || c->rank_h.mode != 0 ||
      c->rank_h.Ns != X->A.Ns || !L->all_fast || X->A.nvec != 1) {
    qsx_set_error("split-chain schedule: needs one hard-core-boson particle sector without symmetry maps "
                  "(N <= 32), bit-parallel terms and one vector");
    return QSX_ERR_UNSUPPORTED;
  }
  const int Np = c->rank_h.Np;
  const int esz = (X->dtype == QSX_F32) ? 4 : (X->dtype == QSX_F64 || X->dtype == QSX_C64) ? 8 : 16;
  // lower half as wide as the LDS budget allows (largest block C(K, c) elements)
  int K = N / 2;
  auto binom = [](int n, int k) -> int64_t {
    if (k < 0 || k > n) return 0;
    int64_t v = 1;
    for (int i = 1; i <= k; ++i) v = v * (n - k + i) / i;
    return v;
  };
  auto maxblk = [&](int k) { int64_t m = 0; for (int q = 0; q <= k; ++q) m = std::max(m, binom(k, q)); return m; };
  while (K > 1 && (maxblk(K) * esz > 64 * 1024 || K > 16)) --K;
  const int U = N - K;
  if (K < 1 || U < 1 || U > 20) { qsx_set_error("split-chain schedule: sector too small or too wide"); return QSX_ERR_UNSUPPORTED; }
  // d grouped by popcount, descending inside a group; position of each d
  std::vector<int32_t> cls_off(K + 2, 0);
  for (int q = 0; q <= K; ++q) cls_off[q + 1] = cls_off[q] + (int32_t)binom(K, q);
  std::vector<uint16_t> dlist((size_t)1 << K), dpos((size_t)1 << K);
  {
    std::vector<int32_t> fill(cls_off.begin(), cls_off.end() - 1);
    for (int64_t d = ((int64_t)1 << K) - 1; d >= 0; --d) {
      const int q = __builtin_popcountll((uint64_t)d);
      dpos[d] = (uint16_t)(fill[q] - cls_off[q]);
      dlist[fill[q]++] = (uint16_t)d;
    }
  }
  // first row of every block, u descending
  std::vector<int64_t> ublk((size_t)1 << U, -1);
  std::vector<int32_t> order;
  int64_t row = 0;
  for (int64_t u = ((int64_t)1 << U) - 1; u >= 0; --u) {
    const int q = Np - __builtin_popcountll((uint64_t)u);
    if (q < 0 || q > K) continue;
    ublk[u] = row;
    row += binom(K, q);
    order.push_back((int32_t)u);
  }
  if (row != X->A.Ns) { qsx_set_error("split-chain schedule: layout does not match the basis"); return QSX_ERR_UNSUPPORTED; }
  // largest blocks first, grouped by the popcount of u
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    const int64_t la = binom(K, Np - __builtin_popcount((uint32_t)a)), lb = binom(K, Np - __builtin_popcount((uint32_t)b));
    if (la != lb) return la > lb;
    return __builtin_popcount((uint32_t)a) < __builtin_popcount((uint32_t)b);
  });
  // the two-pass form's order: by class, u descending inside a class, and
  // its column-strip units
  std::vector<int32_t> ord2;
  std::vector<int32_t> ocls(K + 2, 0);
  std::vector<int64_t> uoff(K + 2, 0);
  for (int q = 0; q <= K; ++q) {
    ocls[q] = (int32_t)ord2.size();
    for (int64_t u = ((int64_t)1 << U) - 1; u >= 0; --u)
      if (Np - __builtin_popcountll((uint64_t)u) == q) ord2.push_back((int32_t)u);
    const int64_t cnt = (int64_t)ord2.size() - ocls[q];
    uoff[q + 1] = uoff[q] + ((binom(K, q) + 63) / 64) * ((cnt + SPLIT2_UNIT_U - 1) / SPLIT2_UNIT_U);
  }
  ocls[K + 1] = (int32_t)ord2.size();
  // terms of the gather form this context uses, by the bits they flip and read
  const std::vector<TermDev>& terms = X->transposed ? L->base.off_h : L->off_hc_h;
  std::vector<TermDev> ell, low, up, upu, cross;
  const uint64_t lmask = ((uint64_t)1 << K) - 1;
  for (const TermDev& T : terms) {
    const uint64_t rd = T.req_mask | T.smask;
    if ((T.flip & ~lmask) == 0) ((rd & ~lmask) == 0 ? ell : low).push_back(T);
    else if ((T.flip & lmask) == 0) ((rd & lmask) == 0 ? upu : up).push_back(T);
    else cross.push_back(T);
  }
  if (upu.size() > 64) { qsx_set_error("split-chain schedule: more than 64 upper-half terms"); return QSX_ERR_UNSUPPORTED; }
  // the lower-half terms that read lower bits only: sliced ELL per class,
  // coefficients through a table of distinct values
  const bool oconj = (X->transposed == X->conjugated);
  std::vector<double2> lut(1, make_double2(0.0, 0.0));
  auto lut_id = [&](double2 v) -> int {
    for (size_t i = 1; i < lut.size(); ++i)
      if (lut[i].x == v.x && lut[i].y == v.y) return (int)i;
    lut.push_back(v);
    return (int)lut.size() - 1;
  };
  std::vector<int64_t> ell_off(K + 2, 0);
  std::vector<int32_t> ell_k(K + 1, 0);
  std::vector<std::vector<uint32_t>> rows_e((size_t)1 << K);
  for (int q = 0; q <= K; ++q) {
    const int nq = cls_off[q + 1] - cls_off[q];
    int km = 0;
    for (int pp = 0; pp < nq; ++pp) {
      const uint32_t d = dlist[cls_off[q] + pp];
      auto& row_e = rows_e[cls_off[q] + pp];
      for (const TermDev& T : ell) {
        const uint64_t v = d & T.req_mask;
        if (v != T.req_val && !(T.has2 && v == T.req_val2)) continue;
        const uint32_t d2 = d ^ (uint32_t)T.flip;
        if (__builtin_popcount(d2) != q) continue;
        const double w = (__builtin_popcountll(d & T.smask) & 1) ? -1.0 : 1.0;
        double2 cf = make_double2(T.cr * w, T.ci * w);
        if (oconj) cf.y = -cf.y;
        row_e.push_back((uint32_t)dpos[d2] | ((uint32_t)lut_id(cf) << 16));
      }
      km = std::max<int>(km, (int)row_e.size());
    }
    ell_k[q] = km;
    ell_off[q + 1] = ell_off[q] + (int64_t)km * nq;
  }
  if (lut.size() > 1024) { qsx_set_error("split-chain schedule: more than 1024 distinct coefficients"); return QSX_ERR_UNSUPPORTED; }
  // two-pass form: per u of ord2, its first row and its applying upper-half
  // terms that read upper bits only (partner first row, coefficient)
  std::vector<int64_t> obase(ord2.size());
  std::vector<int32_t> upn(ord2.size(), 0);
  std::vector<std::vector<std::pair<int64_t, double2>>> upl(ord2.size());
  int upw = 1;
  for (size_t i = 0; i < ord2.size(); ++i) {
    const uint32_t u = (uint32_t)ord2[i];
    obase[i] = ublk[u];
    const uint64_t su = (uint64_t)u << K;
    for (const TermDev& T : upu) {
      const uint64_t v = su & T.req_mask;
      if (v != T.req_val && !(T.has2 && v == T.req_val2)) continue;
      const uint32_t u2 = u ^ (uint32_t)(T.flip >> K);
      if (__builtin_popcount(u2) != __builtin_popcount(u)) continue;
      if (ublk[u2] < 0) continue;
      const double w = (__builtin_popcountll(su & T.smask) & 1) ? -1.0 : 1.0;
      double2 cf = make_double2(T.cr * w, T.ci * w);
      if (oconj) cf.y = -cf.y;
      upl[i].push_back({ublk[u2], cf});
    }
    upn[i] = (int32_t)upl[i].size();
    upw = std::max(upw, (int)upl[i].size());
  }
  // row-pass pairs: consecutive blocks of order with one class, while two
  // staged blocks take at most 64 KB of LDS (larger pairs cost more in
  // occupancy than the shared sliced-ELL entries save)
  const bool pairs = 2 * maxblk(K) * esz <= 64 * 1024;
  std::vector<int32_t> rpair;
  for (size_t i = 0; i < order.size();) {
    const int c1 = Np - __builtin_popcount((uint32_t)order[i]);
    if (pairs && i + 1 < order.size() && Np - __builtin_popcount((uint32_t)order[i + 1]) == c1) {
      rpair.push_back(order[i]);
      rpair.push_back(order[i + 1]);
      i += 2;
    } else {
      rpair.push_back(order[i]);
      rpair.push_back(-1);
      i += 1;
    }
  }
  // the up and crossing terms in compact form, and their partner blocks per u
  std::vector<SplitXT> xts;
  for (const std::vector<TermDev>* lst : {&up, &cross})
    for (const TermDev& T : *lst) {
      SplitXT x;
      x.flip = (uint32_t)T.flip; x.rm = (uint32_t)T.req_mask; x.rv = (uint32_t)T.req_val;
      x.rv2 = (uint32_t)T.req_val2; x.sm = (uint32_t)T.smask; x.has2 = T.has2; x.cr = T.cr; x.ci = T.ci;
      xts.push_back(x);
    }
  const int nxt = (int)xts.size();
  std::vector<int64_t> xb(std::max<size_t>(ord2.size() * (size_t)nxt, 1), -1);
  for (size_t i = 0; i < ord2.size(); ++i)
    for (int t = 0; t < nxt; ++t) xb[i * nxt + t] = ublk[(uint32_t)ord2[i] ^ (xts[t].flip >> K)];
  std::vector<int64_t> upb(ord2.size() * (size_t)upw, -1);
  std::vector<double2> upc(ord2.size() * (size_t)upw, make_double2(0.0, 0.0));
  for (size_t i = 0; i < ord2.size(); ++i)
    for (size_t t = 0; t < upl[i].size(); ++t) {
      upb[i * upw + t] = upl[i][t].first;
      upc[i * upw + t] = upl[i][t].second;
    }
  std::vector<uint32_t> ellv((size_t)std::max<int64_t>(ell_off[K + 1], 1), 0u);
  for (int q = 0; q <= K; ++q) {
    const int nq = cls_off[q + 1] - cls_off[q];
    for (int pp = 0; pp < nq; ++pp) {
      const auto& row_e = rows_e[cls_off[q] + pp];
      for (int k = 0; k < ell_k[q]; ++k)
        ellv[ell_off[q] + (int64_t)k * nq + pp] = k < (int)row_e.size() ? row_e[k] : (uint32_t)pp;
    }
  }
  split_free(X);
  auto up_copy = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
    *dst = nullptr;
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return e;
    X->split_bufs.push_back(*dst);
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  SplitDev H;
  std::memset(&H, 0, sizeof(H));
  H.K = K;
  H.nblk = (int)order.size();
  H.n_low = (int)low.size(); H.n_up = (int)up.size(); H.n_cross = (int)cross.size(); H.n_upu = (int)upu.size();
  H.n_lut = (int)lut.size();
  void* p = nullptr;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) { e = up_copy(low.data(), sizeof(TermDev) * low.size(), &p); H.low = (const TermDev*)p; }
  if (e == hipSuccess) { e = up_copy(up.data(), sizeof(TermDev) * up.size(), &p); H.up = (const TermDev*)p; }
  if (e == hipSuccess) { e = up_copy(cross.data(), sizeof(TermDev) * cross.size(), &p); H.cross = (const TermDev*)p; }
  if (e == hipSuccess) { e = up_copy(upu.data(), sizeof(TermDev) * upu.size(), &p); H.upu = (const TermDev*)p; }
  if (e == hipSuccess) { e = up_copy(ell_off.data(), sizeof(int64_t) * ell_off.size(), &p); H.ell_off = (const int64_t*)p; }
  if (e == hipSuccess) { e = up_copy(ell_k.data(), sizeof(int32_t) * ell_k.size(), &p); H.ell_k = (const int32_t*)p; }
  if (e == hipSuccess) { e = up_copy(ellv.data(), sizeof(uint32_t) * ellv.size(), &p); H.ell = (const uint32_t*)p; }
  if (e == hipSuccess) { e = up_copy(lut.data(), sizeof(double2) * lut.size(), &p); H.lut = (const double2*)p; }
  if (e == hipSuccess) { e = up_copy(dpos.data(), sizeof(uint16_t) * dpos.size(), &p); H.dpos = (const uint16_t*)p; }
  if (e == hipSuccess) { e = up_copy(dlist.data(), sizeof(uint16_t) * dlist.size(), &p); H.dlist = (const uint16_t*)p; }
  if (e == hipSuccess) { e = up_copy(cls_off.data(), sizeof(int32_t) * cls_off.size(), &p); H.cls_off = (const int32_t*)p; }
  if (e == hipSuccess) { e = up_copy(ublk.data(), sizeof(int64_t) * ublk.size(), &p); H.ublk = (const int64_t*)p; }
  if (e == hipSuccess) { e = up_copy(order.data(), sizeof(int32_t) * order.size(), &p); H.order = (const int32_t*)p; }
  if (e == hipSuccess) { e = up_copy(ord2.data(), sizeof(int32_t) * ord2.size(), &p); H.ord2 = (const int32_t*)p; }
  if (e == hipSuccess) { e = up_copy(ocls.data(), sizeof(int32_t) * ocls.size(), &p); H.ocls = (const int32_t*)p; }
  if (e == hipSuccess) { e = up_copy(uoff.data(), sizeof(int64_t) * uoff.size(), &p); H.uoff = (const int64_t*)p; }
  H.nunits = uoff[K + 1];
  H.upw = upw;
  if (e == hipSuccess) { e = up_copy(obase.data(), sizeof(int64_t) * obase.size(), &p); H.obase = (const int64_t*)p; }
  if (e == hipSuccess) { e = up_copy(upn.data(), sizeof(int32_t) * upn.size(), &p); H.upn = (const int32_t*)p; }
  if (e == hipSuccess) { e = up_copy(upb.data(), sizeof(int64_t) * upb.size(), &p); H.upb = (const int64_t*)p; }
  if (e == hipSuccess) { e = up_copy(upc.data(), sizeof(double2) * upc.size(), &p); H.upc = (const double2*)p; }
  if (e == hipSuccess && nxt > 0) { e = up_copy(xts.data(), sizeof(SplitXT) * xts.size(), &p); H.xt = (const SplitXT*)p; }
  if (e == hipSuccess) { e = up_copy(xb.data(), sizeof(int64_t) * xb.size(), &p); H.xb = (const int64_t*)p; }
  H.nrp = (int)(rpair.size() / 2);
  if (e == hipSuccess) { e = up_copy(rpair.data(), sizeof(int32_t) * rpair.size(), &p); H.rpair = (const int32_t*)p; }
  if (e == hipSuccess && with_z) {
    e = hipMalloc(&H.z, (size_t)X->A.Ns * esz);
    if (e == hipSuccess) X->split_bufs.push_back(H.z);
  }
  if (e == hipSuccess) { e = up_copy(&H, sizeof(H), &p); X->split_d = (SplitDev*)p; }
  if (e != hipSuccess) {
    split_free(X);
    qsx_set_error(std::string("split-chain schedule: ") + hipGetErrorString(e));
    return QSX_ERR_HIP;
  }
  X->A.split = X->split_d;
  X->A.split_lds = (int)(maxblk(K) * esz);
  X->A.split_nblk = H.nblk;
  X->A.split_nunits = H.nunits;
  X->A.split_nrp = H.nrp;
  X->A.split_lds2 = (int)((pairs ? 2 : 1) * maxblk(K) * esz);
  X->A.split_pairs = pairs ? 1 : 0;
  X->split_z = with_z;
  X->split_nxt = nxt;
  return QSX_OK;
}

uint64_t qsx_hv_ctx_owner(const qsx_hv_ctx* X) { return X ? X->owner : 0; }

// A changed: captures recorded with the old fields are dropped
static void ctx_changed(qsx_hv_ctx* X) {
  qsx_graph_drop_owner(X->owner);
  X->owner = qsx_owner_new();
}

// Use a sector lookup table (qsx_table_build) instead of the orbit search.
extern "C" int qsx_hv_ctx_set_table(qsx_hv_ctx* X, const void* table, int entry_bytes, int idx_bits, const void* lut,
                                    int lut_len) {
  if (!X || (table && entry_bytes != 4 && entry_bytes != 8) || lut_len < 0 || (lut && lut_len == 0)) {
    qsx_set_error("qsx_hv_ctx_set_table: bad arguments");
    return QSX_ERR_ARG;
  }
  X->A.table = table;
  X->A.table_u64 = entry_bytes == 8;
  X->A.idx_bits = idx_bits;
  X->A.tlut = table ? (const double2*)lut : nullptr;
  X->A.lut_len = X->A.tlut ? lut_len : 0;
  ctx_changed(X);
  return QSX_OK;
}

// 0: one lane per row; 1: compacted work queue; 2/3/4: 2/4/8 lanes per row;
// 5/6: batches of 4/8 terms with rank-based lookups (table or no symmetry);
// 7: split-chain blocks (one hard-core-boson sector without symmetry maps;
// the plan is built on the first request and owned by the context);
// 8/9: direct-rank rows with the terms in chunks of 4/8, gathers predicated
// (a sector without symmetry maps; other products use schedule 0);
// 10: batches of 14 terms where the fast-row kernels apply (as 6 elsewhere);
// 12: as 10, with the workgroups of each XCD on one contiguous run of rows;
// 11: the split-chain layout in two passes, a row pass over the lower-half
// terms into a buffer of Ns elements owned by the context, then a
// column-strip pass over the upper-half and crossing terms (sectors as 7).
// Single vector; several vectors always use one lane per row.
extern "C" int qsx_hv_ctx_set_schedule(qsx_hv_ctx* X, int sched) {
  if (!X || sched < 0 || sched > 12) { qsx_set_error("qsx_hv_ctx_set_schedule: bad arguments"); return QSX_ERR_ARG; }
  if ((sched == 7 && !X->split_d) || (sched == 11 && !X->split_z)) {
    const int rc = split_build(X, sched == 11);
    if (rc) return rc;
  }
  if (sched == 11 && X->split_nxt > SPLIT2_MAX_XT) {
    qsx_set_error("split-chain two-pass schedule: more than 64 terms flip upper bits and read or flip lower bits");
    return QSX_ERR_UNSUPPORTED;
  }
  X->A.sched = sched;
  X->A.xcd_rows = sched == 12 ? 1 : 0;
  ctx_changed(X);
  return QSX_OK;
}

extern "C" int qsx_hv_ctx_create(qsx_core* c, qsx_oplist* l, const void* basis, const void* nrm, int n_bytes,
                                 int64_t Ns, const int64_t* begin, const int64_t* end, int N_p, int dtype,
                                 int64_t nvec, int transposed, int conjugated, int diag_conj, const void* diag_arr,
                                 int diag_arr_cplx, qsx_hv_ctx** out) {
  if (!c || !l || !out || nvec < 1) { qsx_set_error("qsx_hv_ctx_create: bad arguments"); return QSX_ERR_ARG; }
  auto* L = reinterpret_cast<qsx_oplist_full*>(l);
  const bool real = (dtype == QSX_F32 || dtype == QSX_F64);
  if (real && (!L->real_ok || !core_chars_real(c) || (diag_arr && diag_arr_cplx))) {
    qsx_set_error("real dtype requested for an operator with complex matrix elements");
    return QSX_ERR_COMPLEX;
  }
  auto* X = new qsx_hv_ctx();
  std::memset(&X->A, 0, sizeof(HvArgs));
  HvArgs& A = X->A;
  A.C = c->d; A.ph = c->phase; A.R = c->rank_d;
  A.dg = L->base.diag; A.n_diag = L->base.n_diag; A.diag_conj = diag_conj ? 1 : 0;
  A.diag_arr = diag_arr; A.diag_arr_cplx = diag_arr_cplx;
  A.basis = basis; A.nrm = nrm; A.n_bytes = n_bytes; A.begin = begin; A.end = end; A.N_p = N_p;
  A.Ns = Ns; A.nvec = nvec;
  A.alpha_off = make_double2(1, 0); A.alpha_diag = make_double2(1, 0);
  X->c = c;
  X->L = L;
  X->dtype = dtype;
  X->transposed = transposed ? 1 : 0;
  X->conjugated = conjugated ? 1 : 0;
  X->owner = qsx_owner_new();
  *out = X;
  return QSX_OK;
}

// Replace the whole diagonal of the context (its diagonal terms, with their
// conjugation, and any extra diagonal) by the per-row array diag_arr (Ns
// doubles, or Ns (re, im) pairs when diag_arr_cplx).  The caller computes it
// once (the product with a_off = 0 on a vector of ones) and keeps it alive.
extern "C" int qsx_hv_ctx_set_diag(qsx_hv_ctx* X, const void* diag_arr, int diag_arr_cplx) {
  // an empty sector has no diagonal to hold (its array may be a null pointer)
  if (!X || (!diag_arr && X->A.Ns > 0)) { qsx_set_error("qsx_hv_ctx_set_diag: bad arguments"); return QSX_ERR_ARG; }
  if ((X->dtype == QSX_F32 || X->dtype == QSX_F64) && diag_arr_cplx) {
    qsx_set_error("real dtype requested for an operator with a complex diagonal");
    return QSX_ERR_COMPLEX;
  }
  X->A.dg = nullptr;
  X->A.n_diag = 0;
  X->A.diag_conj = 0;
  X->A.diag_arr = diag_arr;
  X->A.diag_arr_cplx = diag_arr_cplx ? 1 : 0;
  ctx_changed(X);
  return QSX_OK;
}

// The captures holding the context's fields go with it.
extern "C" void qsx_hv_ctx_destroy(qsx_hv_ctx* X) {
  if (!X) return;
  qsx_graph_drop_owner(X->owner);
  split_free(X);
  delete X;
}

// Column sums colabs[j] = sum_i |A_ij + shift delta_ij| and the diagonal
// dg[j] = A_jj (Ns (re, im) pairs) of the context's operator A (its
// off-diagonal terms with the context's flags, its diagonal terms and extra
// diagonal).  The inputs of ExpmMultiplyParallel's mu = tr(A)/N and
// ||A - mu I||_1 (_wrapper_csr_trace / _wrapper_csr_1_norm, EMP.html:196-202)
// without building A.  QSX_ERR_UNSUPPORTED when a column holds more than 64
// distinct off-diagonal rows, or for a transposed context whose list does not
// commute with the group; the caller then assembles the CSR instead.
extern "C" int qsx_hv_ctx_column_norms(qsx_hv_ctx* X, double shift_re, double shift_im, double* colabs, double* dg,
                                       void* stream) {
  if (!X || (X->A.Ns > 0 && (!colabs || !dg))) { qsx_set_error("qsx_hv_ctx_column_norms: bad arguments"); return QSX_ERR_ARG; }
  if (X->A.Ns == 0) return QSX_OK;
  const qsx_core* c = X->c;
  const qsx_oplist_full* L = X->L;
  // columns of A = rows of A^T: the flags of the transpose
  const int t2 = X->transposed ? 0 : 1;
  if (c->h.nmaps > 0 && !L->symmetric && !t2) {
    qsx_set_error("column norms of a transposed list that does not commute with the symmetry group");
    return QSX_ERR_UNSUPPORTED;
  }
  hipStream_t st = (hipStream_t)stream;
  HvArgs A = X->A;
  A.off = t2 ? L->base.off : L->off_hc;
  A.n_off = L->base.n_off;
  A.off_conj = (t2 == X->conjugated) ? 1 : 0;
  A.alpha_off = make_double2(1.0, 0.0);
  A.alpha_diag = make_double2(1.0, 0.0);
  A.shift = make_double2(0.0, 0.0);
  A.table = nullptr;
  int* ovf = nullptr;
  QSX_CHECK_HIP(hipMallocAsync((void**)&ovf, sizeof(int), st));
  QSX_CHECK_HIP(hipMemsetAsync(ovf, 0, sizeof(int), st));
  const int lk = pick_lookup(c, A.Ns);
  const double2 sh = make_double2(shift_re, shift_im);
  if (c->h.N <= 32) norm_cols_dispatch<uint32_t>(A, sh, colabs, (double2*)dg, ovf, c->h.nmaps, c->h.fermion != 0, lk, st);
  else norm_cols_dispatch<uint64_t>(A, sh, colabs, (double2*)dg, ovf, c->h.nmaps, c->h.fermion != 0, lk, st);
  QSX_LAUNCH_CHECK();
  int h = 0;
  QSX_CHECK_HIP(hipMemcpyAsync(&h, ovf, sizeof(int), hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipFreeAsync(ovf, st));
  QSX_CHECK_HIP(hipStreamSynchronize(st));
  if (h) {
    qsx_set_error("qsx_hv_ctx_column_norms: a column holds more than 64 distinct rows");
    return QSX_ERR_UNSUPPORTED;
  }
  return QSX_OK;
}

// y = beta*y + a_diag*D x + a_off*Op x
extern "C" int qsx_hv_ctx_apply(qsx_hv_ctx* X, const void* x, void* y, double a_off_re, double a_off_im,
                                double a_diag_re, double a_diag_im, int beta, void* stream) {
  if (!X) { qsx_set_error("qsx_hv_ctx_apply: null context"); return QSX_ERR_ARG; }
  if (X->A.Ns == 0) return QSX_OK;
  const bool real = (X->dtype == QSX_F32 || X->dtype == QSX_F64);
  if (real && (a_off_im != 0.0 || a_diag_im != 0.0)) { qsx_set_error("real vectors with a complex scale"); return QSX_ERR_COMPLEX; }
  HvArgs A = X->A;
  A.v_in = x; A.v_out = y;
  A.alpha_off = make_double2(a_off_re, a_off_im);
  A.alpha_diag = make_double2(a_diag_re, a_diag_im);
  A.beta = beta ? 1 : 0;
  return hv_launch(X->c, X->L, A, X->transposed, X->conjugated, X->dtype, (hipStream_t)stream);
}

// Kernel arguments of a context for the batched wavefront kernels: the
// gather-form term list and the lookup mode.  Sectors whose product needs the
// scatter form, or whose lookup is the orbit search, are not batched.
int qsx_hv_ctx_export(const qsx_hv_ctx* X, HvArgs* A, int* dtype, int* N, int* lookup, int* fermion) {
  if (!X || !A) { qsx_set_error("qsx_hv_batch: null context"); return QSX_ERR_ARG; }
  const qsx_core* c = X->c;
  const qsx_oplist_full* L = X->L;
  *A = X->A;
  const int lk = A->table ? LOOKUP_TABLE : pick_lookup(c, A->Ns);
  if (lk == LOOKUP_SEARCH) {
    qsx_set_error("qsx_hv_batch: every sector needs a lookup table (or a symmetry-free single particle sector)");
    return QSX_ERR_UNSUPPORTED;
  }
  if (c->h.nmaps > 0 && !L->symmetric && !X->transposed && L->base.n_off > 0) {
    qsx_set_error("qsx_hv_batch: operator lists that do not commute with the symmetries are not batched");
    return QSX_ERR_UNSUPPORTED;
  }
  A->off = X->transposed ? L->base.off : L->off_hc;
  A->n_off = L->base.n_off;
  A->off_conj = (X->transposed == X->conjugated) ? 1 : 0;
  A->v_in = nullptr;
  A->v_out = nullptr;
  A->skip = nullptr;
  *dtype = X->dtype;
  *N = c->h.N;
  *lookup = lk;
  *fermion = c->h.fermion != 0;
  return QSX_OK;
}

int qsx_hv_expm_term(void* ctx, const void* x, void* y, double2 coef, double2 shift, const int* skip,
                     hipStream_t st) {
  qsx_hv_ctx* X = (qsx_hv_ctx*)ctx;
  HvArgs A = X->A;
  A.v_in = x; A.v_out = y;
  A.alpha_off = coef; A.alpha_diag = coef; A.shift = shift; A.skip = skip; A.beta = 0;
  return hv_launch(X->c, X->L, A, X->transposed, X->conjugated, X->dtype, st);
}

// The same term with the Taylor epilogue (F += the stored term, the norms of
// the term and of F into nB, nF) fused into the product.  Only the two-pass
// split schedule fuses it; otherwise QSX_ERR_UNSUPPORTED is returned before
// anything is launched, and the caller runs the separate epilogue.
int qsx_hv_expm_term_epi(void* ctx, const void* x, void* y, double2 coef, double2 shift, const int* skip, void* F,
                         double* nB, double* nF, hipStream_t st) {
  qsx_hv_ctx* X = (qsx_hv_ctx*)ctx;
  if (!X->split_z || X->A.sched != 11 || X->A.nvec != 1) return QSX_ERR_UNSUPPORTED;
  HvArgs A = X->A;
  A.v_in = x; A.v_out = y;
  A.alpha_off = coef; A.alpha_diag = coef; A.shift = shift; A.skip = skip; A.beta = 0;
  A.epi_F = F; A.epi_nB = nB; A.epi_nF = nF;
  return hv_launch(X->c, X->L, A, X->transposed, X->conjugated, X->dtype, st);
}

// ---------------------------------------------------------------------------
// sector rank / basis index of arbitrary states (LOOKUP_RANK path), used by
// tests and by callers that index states without a search
// ---------------------------------------------------------------------------
template <typename I>
__global__ void k_sector_rank(const RankDev* __restrict__ R, const I* __restrict__ states, int64_t n,
                              int64_t* __restrict__ out) {
  __shared__ int64_t sb[QSX_SB];
  stage_binom<LOOKUP_RANK>(R, sb);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = sector_rank<I>(R, sb, states[i]);
}

// out[i] = ascending rank of states[i] in the particle sector of the last
// make_basis (-1 outside it); info (8 int64) = {mode, Np, Nup, Ndn, Ns, sector_ok, rank_ok, N}.
extern "C" int qsx_sector_rank(qsx_core* c, const void* states, int64_t n, int64_t* out, int64_t* info,
                               void* stream) {
  if (!c) { qsx_set_error("qsx_sector_rank: null core"); return QSX_ERR_ARG; }
  if (info) {
    info[0] = c->rank_h.mode; info[1] = c->rank_h.Np; info[2] = c->rank_h.Nup; info[3] = c->rank_h.Ndn;
    info[4] = c->rank_h.Ns; info[5] = c->sector_ok; info[6] = c->rank_ok; info[7] = c->h.N;
  }
  if (n <= 0 || !states) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)qsx_grid(n, 256)), blk(256);
  if (c->h.N <= 32) hipLaunchKernelGGL(k_sector_rank<uint32_t>, grid, blk, 0, st, c->rank_d, (const uint32_t*)states, n, out);
  else hipLaunchKernelGGL(k_sector_rank<uint64_t>, grid, blk, 0, st, c->rank_d, (const uint64_t*)states, n, out);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}
