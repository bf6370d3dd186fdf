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
#include <type_traits>
#include "qsx_host.h"

template <typename VT> struct VTr;
// ld/st convert between the stored type and FP64; rnd gives the value that a
// store followed by a load would return (rounding through the stored type),
// without the memory round trip.
template <> struct VTr<float> {
  static constexpr bool cplx = false;
  __device__ static double2 ld(const float* p) { return make_double2((double)*p, 0.0); }
  __device__ static void st(float* p, double2 v) { *p = (float)v.x; }
  __device__ static double2 rnd(double2 v) { return make_double2((double)(float)v.x, 0.0); }
};
template <> struct VTr<double> {
  static constexpr bool cplx = false;
  __device__ static double2 ld(const double* p) { return make_double2(*p, 0.0); }
  __device__ static void st(double* p, double2 v) { *p = v.x; }
  __device__ static double2 rnd(double2 v) { return make_double2(v.x, 0.0); }
};
template <> struct VTr<float2> {
  static constexpr bool cplx = true;
  __device__ static double2 ld(const float2* p) { float2 v = *p; return make_double2((double)v.x, (double)v.y); }
  __device__ static void st(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }
  __device__ static double2 rnd(double2 v) { return make_double2((double)(float)v.x, (double)(float)v.y); }
};
template <> struct VTr<double2> {
  static constexpr bool cplx = true;
  __device__ static double2 ld(const double2* p) { return *p; }
  __device__ static void st(double2* p, double2 v) { *p = v; }
  __device__ static double2 rnd(double2 v) { return v; }
};

// Streaming (non-temporal) element loads and stores: data read or written
// once by a pass, marked so that it does not push reused data out of L2.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
template <typename VT> __device__ __forceinline__ double2 ld_stream(const VT* p);
template <> __device__ __forceinline__ double2 ld_stream<float>(const float* p) {
  return make_double2((double)__builtin_nontemporal_load(p), 0.0);
}
template <> __device__ __forceinline__ double2 ld_stream<double>(const double* p) {
  return make_double2(__builtin_nontemporal_load(p), 0.0);
}
template <> __device__ __forceinline__ double2 ld_stream<float2>(const float2* p) {
  const uint64_t b = __builtin_nontemporal_load((const uint64_t*)p);
  return make_double2((double)__uint_as_float((uint32_t)b), (double)__uint_as_float((uint32_t)(b >> 32)));
}
template <> __device__ __forceinline__ double2 ld_stream<double2>(const double2* p) {
  const nt_d2 v = __builtin_nontemporal_load((const nt_d2*)p);
  return make_double2(v.x, v.y);
}
template <typename VT> __device__ __forceinline__ void st_stream(VT* p, double2 v);
template <> __device__ __forceinline__ void st_stream<float>(float* p, double2 v) {
  __builtin_nontemporal_store((float)v.x, p);
}
template <> __device__ __forceinline__ void st_stream<double>(double* p, double2 v) {
  __builtin_nontemporal_store(v.x, p);
}
template <> __device__ __forceinline__ void st_stream<float2>(float2* p, double2 v) {
  const uint64_t b = (uint64_t)__float_as_uint((float)v.x) | ((uint64_t)__float_as_uint((float)v.y) << 32);
  __builtin_nontemporal_store(b, (uint64_t*)p);
}
template <> __device__ __forceinline__ void st_stream<double2>(double2* p, double2 v) {
  nt_d2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, (nt_d2*)p);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// sqrt of a small non-negative integer (stabiliser sizes): exact table for
// n < 64, hardware sqrt above.
__constant__ double c_sqrt_tab[64] = {
    0.0, 1.0, 1.4142135623730951, 1.7320508075688772, 2.0, 2.23606797749979, 2.449489742783178,
    2.6457513110645907, 2.8284271247461903, 3.0, 3.1622776601683795, 3.3166247903554, 3.4641016151377544,
    3.605551275463989, 3.7416573867739413, 3.872983346207417, 4.0, 4.123105625617661, 4.242640687119285,
    4.358898943540674, 4.47213595499958, 4.58257569495584, 4.69041575982343, 4.795831523312719,
    4.898979485566356, 5.0, 5.0990195135927845, 5.196152422706632, 5.291502622129181, 5.385164807134504,
    5.477225575051661, 5.5677643628300215, 5.656854249492381, 5.744562646538029, 5.830951894845301,
    5.916079783099616, 6.0, 6.082762530298219, 6.164414002968976, 6.244997998398398, 6.324555320336759,
    6.4031242374328485, 6.48074069840786, 6.557438524302, 6.6332495807108, 6.708203932499369,
    6.782329983125268, 6.855654600401044, 6.928203230275509, 7.0, 7.0710678118654755, 7.14142842854285,
    7.211102550927978, 7.280109889280518, 7.3484692283495345, 7.416198487095663, 7.483314773547883,
    7.54983443527075, 7.615773105863909, 7.681145747868608, 7.745966692414834, 7.810249675906654,
    7.874007874011811, 7.937253933193772};

__device__ __forceinline__ double sqrt_small(int n) { return n < 64 ? c_sqrt_tab[n] : sqrt((double)n); }

// sum over G consecutive lanes (G divides the wavefront size)
template <int G>
__device__ __forceinline__ double grp_sum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

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
  // with a coefficient table the entry's code field is an id into tlut
  // (sign-free chi(g) * sqrt(n_c) of the target), so no n_c read is needed
  const double2* tlut;
  int lut_len;        // entries of tlut
  int n_sb;           // binomial entries staged by k_hv_rows_fast_batch (even)
  int sched;          // 0: one lane per row, 1: compacted queue (one vector)
  int fast_rows;      // every off-diagonal term bit-parallel, no diagonal terms, single hcb sector
  const struct SplitDev* split;  // split-chain plan (schedule 7), device memory
  int split_lds;      // bytes of LDS the split kernel stages (largest block)
  int split_nblk;     // its workgroups (blocks of the layout)
  int64_t split_nunits;  // column-strip units of the two-pass form (schedule 11)
  int split_nrp;         // its row-pass workgroups (pairs of blocks, or single blocks)
  int split_lds2;        // bytes of LDS its row pass stages
  int split_pairs;       // its row pass takes two blocks per workgroup
  int xcd_rows;          // k_hv_rows_fast_batch: each XCD's workgroups take one contiguous run of rows
  // expm Taylor epilogue fused into the two-pass form's strip pass (one
  // vector): F += the stored term, and the maxima of |term| and |F| go to
  // the two norm slots (null: a plain product)
  void* epi_F;
  double* epi_nB;
  double* epi_nF;
};

// Split-chain layout of one hard-core-boson particle sector of N <= 32 bits
// without symmetry maps: a state is s = (u << K) | d, u the upper N - K
// bits and d the lower K bits.  The descending basis lists every u in
// descending order and, for each, the d of popcount Np - popc(u) in
// descending order: the rows of one u form a contiguous block, a term that
// flips only lower bits maps the block onto itself, and one that flips only
// upper bits maps block u onto block u ^ flip at the same position.
struct SplitDev {
  int K, nblk;
  int n_low, n_up, n_cross, n_upu, n_lut;
  const TermDev* low;     // terms flipping lower bits only, reading some upper bit
  const TermDev* up;      // terms flipping upper bits only, reading some lower bit
  const TermDev* cross;   // terms flipping bits of both halves
  const TermDev* upu;     // terms flipping and reading upper bits only (uniform over a block)
  // the lower-bit terms that read lower bits only, as one sparse matrix per
  // popcount class c of d in sliced-ELL form: entry k of position p at
  // ell_off[c] + k * n_c + p (n_c = C(K, c)), ell_k[c] entries per position,
  // each (partner position) | (coefficient id << 16); padding (p, id 0 = 0)
  const int64_t* ell_off;
  const int32_t* ell_k;
  const uint32_t* ell;
  const double2* lut;     // n_lut coefficients (lut[0] = 0)
  const uint16_t* dpos;   // [2^K] position of d among the d of its popcount, descending
  const uint16_t* dlist;  // [2^K] the d grouped by popcount, descending inside a group
  const int32_t* cls_off; // [K + 2] group offsets into dlist
  const int64_t* ublk;    // [2^(N-K)] first row of the block of u (-1: no such block)
  const int32_t* order;   // [nblk] u of each workgroup (largest blocks first)
  // two-pass form (schedule 11): ord2 lists the u by class c = Np - popc(u)
  // ascending, u descending inside a class, class c at ord2[ocls[c] ..
  // ocls[c + 1]); the column-strip pass has units (class c, strip of 64
  // positions, group of 16 u), uoff[c] .. uoff[c + 1] strip-major, nunits in
  // all; z holds the row pass' partial sums (one vector in the vector dtype)
  const int32_t* ord2;    // [nblk]
  const int32_t* ocls;    // [K + 2]
  const int64_t* uoff;    // [K + 2]
  int64_t nunits;
  void* z;
  // per u of ord2 (index i): its first row, and its upper-half terms that
  // read upper bits only, compacted: upn[i] of them at i * upw .. with the
  // partner block's first row and the coefficient (conjugation applied)
  const int64_t* obase;   // [nblk]
  const int32_t* upn;     // [nblk]
  const int64_t* upb;     // [nblk * upw]
  const double2* upc;     // [nblk * upw]
  int upw;
  // the up and crossing terms (n_up + n_cross <= SPLIT2_MAX_XT) in compact
  // form, and per u of ord2 the first row of each term's partner block
  // (u ^ the term's upper flip; -1: none)
  const struct SplitXT* xt;  // [n_up + n_cross]
  const int64_t* xb;         // [nblk * (n_up + n_cross)]
  // row pass: workgroup b takes the blocks rpair[2b] and rpair[2b + 1] (-1:
  // none), two u of one class, so each sliced-ELL entry serves both
  const int32_t* rpair;      // [2 * nrp]
  int nrp;
};

// A term of the column-strip pass in compact form (N <= 32 bits).
struct SplitXT {
  uint32_t flip, rm, rv, rv2, sm;
  int32_t has2;
  double cr, ci;
};
constexpr int SPLIT2_MAX_XT = 64;

// Block index with the blocks of each XCD (workgroups are dispatched
// round-robin over the 8 XCDs) renumbered into one contiguous run.
__device__ __forceinline__ int64_t xcd_contiguous(int64_t b, int64_t nb) {
  const int64_t x = b & 7, pos = b >> 3, q = nb >> 3, r = nb & 7;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + pos;
}

// Block index with each XCD taking runs of G consecutive blocks inside a
// window of 8 G that all XCDs share (G = 1: dispatch order; one run per XCD:
// xcd_contiguous).  Blocks past the last full window keep their index.
__device__ __forceinline__ int64_t xcd_chunked(int64_t b, int64_t nb, int64_t G) {
  const int64_t nfull = (nb / (8 * G)) * 8 * G;
  if (b >= nfull) return b;
  const int64_t pos = b >> 3;
  return (pos / G) * 8 * G + (b & 7) * G + pos % G;
}

__device__ __forceinline__ double load_n(const void* p, int nb, int64_t i) {
  return nb == 1 ? (double)((const uint8_t*)p)[i] : nb == 2 ? (double)((const uint16_t*)p)[i]
                                                              : (double)((const uint32_t*)p)[i];
}

template <typename I>
__device__ __forceinline__ int64_t sector_rank(const RankDev* __restrict__ R, const int64_t* sb, I s);

template <typename I, int LK>
__device__ __forceinline__ int64_t lookup_index(const HvArgs& A, const int64_t* sb, I rep) {
  if constexpr (LK == LOOKUP_RANK) {
    const int64_t rk = sector_rank<I>(A.R, sb, rep);
    return rk < 0 ? -1 : A.R->Ns - 1 - rk;
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

// Binomial table C(n, k), n <= 64, k <= 32, staged in LDS by the kernels
// that rank states (rows of 33 entries).  Ranks only read rows n < N (bit
// positions), so a block stages rows 0..N and leaves the rest unset.
constexpr int QSX_SB = 65 * 33;

template <int LK>
__device__ __forceinline__ void stage_binom(const RankDev* __restrict__ R, int64_t* sb) {
  if constexpr (LK == LOOKUP_TABLE || LK == LOOKUP_RANK) {
    const int n = min(QSX_SB, (R->N + 1) * 33);
    const int64_t* __restrict__ b = &R->binom[0][0];
    for (int i = threadIdx.x; i < n; i += blockDim.x) sb[i] = b[i];
  }
}

template <typename I>
__device__ __forceinline__ int64_t comb_rank_sb(const int64_t* sb, I s) {
  int64_t r = 0;
  int j = 1;
  while (s) {
    const int p = (sizeof(I) == 8) ? __builtin_ctzll((uint64_t)s) : __builtin_ctz((uint32_t)s);
    s &= s - 1;
    r += sb[p * 33 + j];
    ++j;
  }
  return r;
}

// Ascending rank of s inside the particle sector described by R, -1 outside.
// Single exit on purpose: an earlier multi-return form of this function was
// compiled for gfx950 with the mode-2 result never written (the value of the
// trailing return was lost when the uniform branches were structurised).
template <typename I>
__device__ __forceinline__ int64_t sector_rank(const RankDev* __restrict__ R, const int64_t* sb, I s) {
  const int mode = R->mode;
  int64_t res = (int64_t)s;
  if (mode == 0) {
    res = (popc<I>(s) == R->Np) ? comb_rank_sb<I>(sb, s) : (int64_t)-1;
  } else if (mode == 1) {
    const int L = R->L;
    const I mlo = (I)((((uint64_t)1) << L) - 1);
    const I up = s >> L, dn = s & mlo;
    const bool ok = (popc<I>(up) == R->Nup) && (popc<I>(dn) == R->Ndn);
    res = ok ? comb_rank_sb<I>(sb, up) * R->cdn + comb_rank_sb<I>(sb, dn) : (int64_t)-1;
  }
  return res;
}

// Sum of C(p, j) over the set bits p of x inside [lo, hi], j = number of set
// bits of x at positions <= p.
template <typename I>
__device__ __forceinline__ int64_t window_rank(const int64_t* sb, I x, int lo, int hi) {
  constexpr int W = 8 * (int)sizeof(I);
  I w = (x >> lo) << lo;
  if (hi < W - 1) w &= (((I)2) << hi) - (I)1;
  int j = (lo > 0) ? popc<I>(x & ((((I)1) << lo) - (I)1)) : 0;
  int64_t r = 0;
  while (w) {
    const int p = (sizeof(I) == 8) ? __builtin_ctzll((uint64_t)w) : __builtin_ctz((uint32_t)w);
    w &= w - 1;
    ++j;
    r += sb[p * 33 + j];
  }
  return r;
}

// Rank data of a row state: full rank (mode 0), or up / down ranks (mode 1).
struct RowRank {
  int64_t a, b;
};

template <typename I>
__device__ __forceinline__ RowRank row_rank(const RankDev* __restrict__ R, const int64_t* sb, I s) {
  RowRank rr{0, 0};
  if (R->mode == 0) {
    rr.a = comb_rank_sb<I>(sb, s);
  } else if (R->mode == 1) {
    const int L = R->L;
    const I mlo = (I)((((uint64_t)1) << L) - 1);
    rr.a = comb_rank_sb<I>(sb, s >> L);
    rr.b = comb_rank_sb<I>(sb, s & mlo);
  }
  return rr;
}

// Rank of s2 = s with the bits of [lo, hi] changed, from the rank of s.
// Single exit (see sector_rank).
template <typename I>
__device__ __forceinline__ int64_t sector_rank_inc(const RankDev* __restrict__ R, const int64_t* sb, I s, RowRank rr,
                                                   I s2, int lo, int hi) {
  const int mode = R->mode;
  int64_t res = (int64_t)s2;
  if (mode == 0) {
    const bool ok = popc<I>(s2) == R->Np;
    res = ok ? rr.a - window_rank<I>(sb, s, lo, hi) + window_rank<I>(sb, s2, lo, hi) : (int64_t)-1;
  } else if (mode == 1) {
    const int L = R->L;
    const I mlo = (I)((((uint64_t)1) << L) - 1);
    const I up = s2 >> L, dn = s2 & mlo;
    const bool ok = (popc<I>(up) == R->Nup) && (popc<I>(dn) == R->Ndn);
    int64_t ra = rr.a, rb = rr.b;
    if (hi < L) {
      rb = rr.b - window_rank<I>(sb, s & mlo, lo, hi) + window_rank<I>(sb, dn, lo, hi);
    } else if (lo >= L) {
      ra = rr.a - window_rank<I>(sb, s >> L, lo - L, hi - L) + window_rank<I>(sb, up, lo - L, hi - L);
    } else {
      ra = comb_rank_sb<I>(sb, up);
      rb = comb_rank_sb<I>(sb, dn);
    }
    res = ok ? ra * R->cdn + rb : (int64_t)-1;
  }
  return res;
}

// Element factor of a target reached through group element `code` with
// stabiliser count cnt and fermion sign: sign * chi(g) * sqrt(n_c).
__device__ __forceinline__ double2 elem_fac(const HvArgs& A, int code, int cnt, int sign) {
  const double w = (double)sign * sqrt_small(cnt);
  const double2 p = A.ph[code];
  return make_double2(w * p.x, w * p.y);
}

// Basis index of the representative of the sector state of rank rk, and its
// element factor.  Returns false when the representative is not in the basis.
template <typename I>
__device__ __forceinline__ bool table_entry(const HvArgs& A, int64_t rk, int64_t& c, double2& fac) {
  if (rk < 0) return false;
  int code, sign = 1;
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
  if (A.tlut) {
    const double2 f = A.tlut[code];
    fac = make_double2(sign * f.x, sign * f.y);
  } else {
    fac = elem_fac(A, code, (int)load_n(A.nrm, A.n_bytes, c), sign);
  }
  return true;
}

// Basis index c of the representative of s2 and the element factor (1 with
// no symmetry maps).  Returns false when s2 has no representative in the basis.
template <typename I, int M, bool FERM, int LK>
__device__ __forceinline__ bool resolve(const HvArgs& A, const int64_t* sb, I s2, int64_t& c, double2& fac) {
  fac = make_double2(1.0, 0.0);
  if constexpr (LK == LOOKUP_TABLE) {
    return table_entry<I>(A, sector_rank<I>(A.R, sb, s2), c, fac);
  } else {
    I rep = s2;
    if constexpr (M != 0) {
      OrbitRes R;
      rep = orbit_max<I, M, FERM>(A.C, s2, R);
      fac = elem_fac(A, R.code, R.cnt, R.sign);
    }
    c = lookup_index<I, LK>(A, sb, rep);
    return c >= 0;
  }
}

// As resolve, with the rank of s2 obtained incrementally from the row state s
// (rank data rr) when the lookup works by rank.
template <typename I, int M, bool FERM, int LK>
__device__ __forceinline__ bool resolve_row(const HvArgs& A, const int64_t* sb, I s, RowRank rr,
                                            const TermDev* __restrict__ T, I s2, int64_t& c, double2& fac) {
  if constexpr (LK == LOOKUP_TABLE) {
    fac = make_double2(1.0, 0.0);
    return table_entry<I>(A, sector_rank_inc<I>(A.R, sb, s, rr, s2, T->lo, T->hi), c, fac);
  } else if constexpr (LK == LOOKUP_RANK) {
    fac = make_double2(1.0, 0.0);
    const int64_t rk = sector_rank_inc<I>(A.R, sb, s, rr, s2, T->lo, T->hi);
    c = rk < 0 ? (int64_t)-1 : A.R->Ns - 1 - rk;
    return c >= 0;
  } else {
    return resolve<I, M, FERM, LK>(A, sb, s2, c, fac);
  }
}

template <typename I, int LK>
__device__ __forceinline__ RowRank row_rank_if(const HvArgs& A, const int64_t* sb, I s) {
  if constexpr (LK == LOOKUP_TABLE || LK == LOOKUP_RANK) return row_rank<I>(A.R, sb, s);
  else return RowRank{0, 0};
}

// Reduced element of term T applied to state s of a row with normalisation
// 1/sqrt(n_r) = rsq.  Calls f(c, coef) when non-zero.  CPLX=false assumes the
// host checked that every factor is real.
template <typename I, int M, bool FERM, int LK, bool CPLX, typename F>
__device__ __forceinline__ void term_element(const HvArgs& A, const int64_t* sb, RowRank rr,
                                             const TermDev* __restrict__ T, I s, double rsq, int conj, F&& f) {
  I s2 = s;
  double mag;
  int ip;
  if (!apply_term<I>(T, A.C->fermion != 0, s2, mag, ip)) return;
  int64_t c;
  double2 fac;
  if (!resolve_row<I, M, FERM, LK>(A, sb, s, rr, T, s2, c, fac)) return;
  double w = mag;
  if constexpr (M != 0) w *= rsq;
  if constexpr (CPLX) {
    double2 coef = cmul(make_double2(T->jr, T->ji), ipow_c(ip));
    coef = make_double2(coef.x * w, coef.y * w);
    if constexpr (M != 0) coef = cmul(coef, fac);
    if (conj) coef = cconj(coef);
    f(c, coef);
  } else {
    double v = T->jr * w * ((ip & 2) ? -1.0 : 1.0);
    if constexpr (M != 0) v *= fac.x;
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
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  if (A.skip && *A.skip) return;
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const I s = ((const I*)A.basis)[r];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
  const RowRank rr = row_rank_if<I, LK>(A, sb, s);
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const double2 dsum = row_diag<I, M, FERM, LK, CPLX>(A, r, s);
  if constexpr (NV == 1) {
    double2 acc = make_double2(0.0, 0.0);
    for (int t = 0; t < A.n_off; ++t) {
      term_element<I, M, FERM, LK, CPLX>(A, sb, rr, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
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
      term_element<I, M, FERM, LK, CPLX>(A, sb, rr, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
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

// One vector, one lane per row, for the common case the host flags with
// fast_rows: every off-diagonal term is in the bit-parallel form, the
// diagonal is a folded per-row array (no diagonal terms) and the lookup is
// the combinadic rank of one hard-core-boson sector (through the sector
// table, or directly without symmetry).  Same elements as k_hv_rows without
// the generic operator-string and multi-mode rank paths in the loop.
// RPL rows per lane (r, r + 256, ... within the block's 256 RPL rows): the
// uniform term fields are read once for all of them.  One term at a time:
// a wave skips a term no lane of it needs, and a lane skips the loads of a
// term that gives its row nothing.
template <typename I, typename VT, int M, int LK, int RPL>
__global__ void __launch_bounds__(256) k_hv_rows_fast(HvArgs A) {
  static_assert(LK == LOOKUP_TABLE || LK == LOOKUP_RANK, "rank-based lookups only");
  __shared__ int64_t sb[QSX_SB];
  if (A.skip && *A.skip) return;
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x * RPL + threadIdx.x;
  if (r0 >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const int Np = A.R->Np;
  int64_t rr[RPL];
  bool live[RPL];
  I s[RPL];
  int64_t ra[RPL];
  double rsq[RPL];
  double2 acc[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    rr[k] = r0 + (int64_t)k * blockDim.x;
    live[k] = rr[k] < A.Ns;
    const int64_t rc = live[k] ? rr[k] : r0;
    s[k] = ((const I*)A.basis)[rc];
    rsq[k] = 1.0;
    if constexpr (M != 0) rsq[k] = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, rc));
    ra[k] = comb_rank_sb<I>(sb, s[k]);
    acc[k] = make_double2(0.0, 0.0);
  }
  for (int t = 0; t < A.n_off; ++t) {
    const TermDev* __restrict__ T = A.off + t;
    const I rq_m = (I)T->req_mask, rq_v = (I)T->req_val, rq_v2 = (I)T->req_val2, fl = (I)T->flip;
    const bool has2 = T->has2 != 0;
    const int lo = T->lo, hi = T->hi;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const I sk = s[k];
      const I v = sk & rq_m;
      if (v != rq_v && !(has2 && v == rq_v2)) continue;
      const I s2 = sk ^ fl;
      if (popc<I>(s2) != Np) continue;
      int64_t rk;
      if (hi == lo + 1) {
        // one particle hops between adjacent bits lo and lo + 1: with j
        // particles below lo, its term of the rank changes by
        // C(lo + 1, j + 1) - C(lo, j + 1) = C(lo, j)
        const int64_t dl = sb[lo * 33 + popc<I>(sk & ((((I)1) << lo) - (I)1))];
        rk = ((sk >> lo) & (I)1) ? ra[k] + dl : ra[k] - dl;
      } else {
        rk = ra[k] - window_rank<I>(sb, sk, lo, hi) + window_rank<I>(sb, s2, lo, hi);
      }
      int64_t c;
      double2 fac = make_double2(1.0, 0.0);
      double2 x;
      if constexpr (LK == LOOKUP_TABLE) {
        // decode the entry, then issue the x gather and the coefficient
        // read together: both depend only on the entry
        int code;
        double sg = 1.0;
        if (A.table_u64) {
          const uint64_t e = ((const uint64_t*)A.table)[rk];
          if (e == ~0ull) continue;
          c = (int64_t)(e & ((1ull << 40) - 1));
          code = (int)((e >> 40) & 0x7FFFFF);
          if (e >> 63) sg = -1.0;
        } else {
          const uint32_t e = ((const uint32_t*)A.table)[rk];
          if (e == 0xFFFFFFFFu) continue;
          c = (int64_t)(e & ((1u << A.idx_bits) - 1u));
          code = (int)(e >> A.idx_bits);
        }
        x = VTr<VT>::ld(vin + c);
        const double2 f = A.tlut[code];
        fac = make_double2(sg * f.x, sg * f.y);
      } else {
        c = A.R->Ns - 1 - rk;
        x = VTr<VT>::ld(vin + c);
      }
      double w = (popc<I>(sk & (I)T->smask) & 1) ? -1.0 : 1.0;
      if constexpr (M != 0) w *= rsq[k];
      if constexpr (CPLX) {
        double2 coef = make_double2(T->cr * w, T->ci * w);
        if constexpr (M != 0) coef = cmul(coef, fac);
        if (A.off_conj) coef = cconj(coef);
        acc[k] = cadd(acc[k], cmul(coef, x));
      } else {
        double vv = T->cr * w;
        if constexpr (M != 0) vv *= fac.x;
        acc[k].x += vv * x.x;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    if (!live[k]) continue;
    const int64_t r = rr[k];
    double2 d = make_double2(0.0, 0.0);
    if (A.diag_arr) {
      if (A.diag_arr_cplx) d = ((const double2*)A.diag_arr)[r];
      else d.x = ((const double*)A.diag_arr)[r];
    }
    const double2 dsum = cadd(cmul(d, A.alpha_diag), A.shift);
    const double2 xr = VTr<VT>::ld(vin + r);
    double2 res;
    if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc[k]), cmul(dsum, xr));
    else res = make_double2(A.alpha_off.x * acc[k].x + dsum.x * xr.x, 0.0);
    if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
    VTr<VT>::st(vout + r, res);
  }
}

// k_hv_rows_fast for the direct rank (no table), the terms in chunks of T:
// first the T terms' tests and ranks for the lane's RPL rows, then all the
// gathers of x the chunk's applicable terms need, each under its own
// predicate (no loads for terms that give a row nothing, so nothing is read
// for them), then the products.  A chunk no lane of the wave needs is
// skipped as a whole.
template <typename I, typename VT, int RPL, int T>
__global__ void __launch_bounds__(256) k_hv_rows_fast_chunk(HvArgs A) {
  __shared__ int64_t sb[QSX_SB];
  if (A.skip && *A.skip) return;
  stage_binom<LOOKUP_RANK>(A.R, sb);
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x * RPL + threadIdx.x;
  if (r0 >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const int Np = A.R->Np;
  const int64_t last = A.R->Ns - 1;
  int64_t rr[RPL];
  bool live[RPL];
  I s[RPL];
  int64_t ra[RPL];
  double2 acc[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    rr[k] = r0 + (int64_t)k * blockDim.x;
    live[k] = rr[k] < A.Ns;
    s[k] = ((const I*)A.basis)[live[k] ? rr[k] : r0];
    ra[k] = comb_rank_sb<I>(sb, s[k]);
    acc[k] = make_double2(0.0, 0.0);
  }
  const int nt = A.n_off;
  for (int t0 = 0; t0 < nt; t0 += T) {
    bool ok[T][RPL];
    int64_t c[T][RPL];
    bool any = false;
#pragma unroll
    for (int b = 0; b < T; ++b) {
      const bool tv = t0 + b < nt;
      const TermDev* __restrict__ Tm = A.off + (tv ? t0 + b : nt - 1);
      const I rq_m = (I)Tm->req_mask, rq_v = (I)Tm->req_val, rq_v2 = (I)Tm->req_val2, fl = (I)Tm->flip;
      const bool has2 = Tm->has2 != 0;
      const int lo = Tm->lo, hi = Tm->hi;
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const I sk = s[k];
        const I v = sk & rq_m;
        const I s2 = sk ^ fl;
        const bool a = tv && live[k] && (v == rq_v || (has2 && v == rq_v2)) && popc<I>(s2) == Np;
        int64_t q = 0;
        if (a) {
          if (hi == lo + 1) {
            const int64_t dl = sb[lo * 33 + popc<I>(sk & ((((I)1) << lo) - (I)1))];
            q = ((sk >> lo) & (I)1) ? ra[k] + dl : ra[k] - dl;
          } else {
            q = ra[k] - window_rank<I>(sb, sk, lo, hi) + window_rank<I>(sb, s2, lo, hi);
          }
        }
        ok[b][k] = a;
        c[b][k] = last - q;
        any = any || a;
      }
    }
    if (!__any(any)) continue;
    double2 x[T][RPL];
#pragma unroll
    for (int b = 0; b < T; ++b)
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        x[b][k] = make_double2(0.0, 0.0);
        if (ok[b][k]) x[b][k] = VTr<VT>::ld(vin + c[b][k]);
      }
#pragma unroll
    for (int b = 0; b < T; ++b) {
      if (t0 + b >= nt) break;
      const TermDev* __restrict__ Tm = A.off + t0 + b;
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        if (!ok[b][k]) continue;
        const double w = (popc<I>(s[k] & (I)Tm->smask) & 1) ? -1.0 : 1.0;
        if constexpr (CPLX) {
          double2 coef = make_double2(Tm->cr * w, Tm->ci * w);
          if (A.off_conj) coef = cconj(coef);
          acc[k] = cadd(acc[k], cmul(coef, x[b][k]));
        } else {
          acc[k].x += Tm->cr * w * x[b][k].x;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    if (!live[k]) continue;
    const int64_t r = rr[k];
    double2 d = make_double2(0.0, 0.0);
    if (A.diag_arr) {
      if (A.diag_arr_cplx) d = ((const double2*)A.diag_arr)[r];
      else d.x = ((const double*)A.diag_arr)[r];
    }
    const double2 dsum = cadd(cmul(d, A.alpha_diag), A.shift);
    const double2 xr = VTr<VT>::ld(vin + r);
    double2 res;
    if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc[k]), cmul(dsum, xr));
    else res = make_double2(A.alpha_off.x * acc[k].x + dsum.x * xr.x, 0.0);
    if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
    VTr<VT>::st(vout + r, res);
  }
}

// Off-diagonal sums of RPL rows (states s, ranks ra, 1/sqrt(n) rsq) over all
// terms of A, the terms in batches of B: first the B ranks, then the B table
// reads (TW-byte entries; TW = 0: direct rank, no table), then the B gathers
// of x (and coefficient reads), then the B products, so a lane keeps B
// independent requests in flight instead of one dependent chain per term.
// Rows a term does not reach read entry 0 of the table and of x and add
// zero: the loads stay unconditional, so none waits behind a branch.  sb:
// the binomials in LDS; slut: the coefficient table in LDS (LUTL).
template <typename I, typename VT, int M, int LK, int RPL, int B, int TW, bool LUTL>
__device__ __forceinline__ void fast_rows_acc(const HvArgs& A, const int64_t* sb, const double2* slut,
                                              const VT* __restrict__ vin, const I* s, const int64_t* ra,
                                              const double* rsq, double2* acc) {
  static_assert(LK == LOOKUP_TABLE || LK == LOOKUP_RANK, "rank-based lookups only");
  static_assert((LK == LOOKUP_TABLE) == (TW == 4 || TW == 8), "table entries are 4 or 8 bytes");
  using E = typename std::conditional<TW == 8, uint64_t, uint32_t>::type;
  constexpr bool CPLX = VTr<VT>::cplx;
  const int Np = A.R->Np;
  const int nt = A.n_off;
  for (int t0 = 0; t0 < nt; t0 += B) {
    // ranks of the images (rank 0, a valid index, where the term gives none)
    int64_t rk[B][RPL];
    bool ok[B][RPL];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const bool tv = t0 + b < nt;
      const TermDev* __restrict__ T = A.off + (tv ? t0 + b : nt - 1);
      const I rq_m = (I)T->req_mask, rq_v = (I)T->req_val, rq_v2 = (I)T->req_val2, fl = (I)T->flip;
      const bool has2 = T->has2 != 0;
      const int lo = T->lo, hi = T->hi;
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const I sk = s[k];
        const I v = sk & rq_m;
        const I s2 = sk ^ fl;
        const bool a = tv && (v == rq_v || (has2 && v == rq_v2)) && popc<I>(s2) == Np;
        int64_t q = 0;
        if (hi == lo + 1) {
          // one particle hops between adjacent bits lo and lo + 1: with j
          // particles below lo, its term of the rank changes by
          // C(lo + 1, j + 1) - C(lo, j + 1) = C(lo, j)
          const int64_t dl = sb[lo * 33 + popc<I>(sk & ((((I)1) << lo) - (I)1))];
          q = ((sk >> lo) & (I)1) ? ra[k] + dl : ra[k] - dl;
        } else if (a) {
          q = ra[k] - window_rank<I>(sb, sk, lo, hi) + window_rank<I>(sb, s2, lo, hi);
        }
        ok[b][k] = a;
        rk[b][k] = a ? q : 0;
      }
    }
    // basis index and element factor of each image, and the gathers
    double2 x[B][RPL], fac[B][RPL];
    if constexpr (LK == LOOKUP_TABLE) {
      E e[B][RPL];
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int k = 0; k < RPL; ++k) e[b][k] = ((const E*)A.table)[rk[b][k]];
#pragma unroll
      for (int b = 0; b < B; ++b) {
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const E ek = e[b][k];
          const bool hit = ok[b][k] && ek != ~(E)0;
          int64_t c;
          int code;
          double sg = 1.0;
          if constexpr (TW == 8) {
            c = (int64_t)(ek & ((1ull << 40) - 1));
            code = (int)((ek >> 40) & 0x7FFFFF);
            if (ek >> 63) sg = -1.0;
          } else {
            c = (int64_t)(ek & ((1u << A.idx_bits) - 1u));
            code = (int)(ek >> A.idx_bits);
          }
          c = hit ? c : 0;
          code = hit ? code : 0;
          ok[b][k] = hit;
          x[b][k] = VTr<VT>::ld(vin + c);
          double2 f;
          if constexpr (LUTL) f = slut[code];
          else f = A.tlut[code];
          fac[b][k] = make_double2(sg * f.x, sg * f.y);
        }
      }
    } else {
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          x[b][k] = VTr<VT>::ld(vin + (A.R->Ns - 1 - rk[b][k]));
          fac[b][k] = make_double2(1.0, 0.0);
        }
    }
    // products (zero weight where the term gives no element)
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const TermDev* __restrict__ T = A.off + (t0 + b < nt ? t0 + b : nt - 1);
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        double w = (popc<I>(s[k] & (I)T->smask) & 1) ? -1.0 : 1.0;
        if constexpr (M != 0) w *= rsq[k];
        w = ok[b][k] ? w : 0.0;
        if constexpr (CPLX) {
          double2 coef = make_double2(T->cr * w, T->ci * w);
          if constexpr (M != 0) coef = cmul(coef, fac[b][k]);
          if (A.off_conj) coef = cconj(coef);
          acc[k] = cadd(acc[k], cmul(coef, x[b][k]));
        } else {
          double vv = T->cr * w;
          if constexpr (M != 0) vv *= fac[b][k].x;
          acc[k].x += vv * x[b][k].x;
        }
      }
    }
  }
}

// Stage the binomials (A.n_sb rows' worth) and, with LUTL, the coefficient
// table into the dynamic LDS of a fast-row kernel; returns the two arrays.
template <bool LUTL>
__device__ __forceinline__ void fast_stage(const HvArgs& A, int64_t* smem, const int64_t*& sb, const double2*& slut) {
  const int n = min(A.n_sb, QSX_SB);
  const int64_t* __restrict__ bsrc = &A.R->binom[0][0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) smem[i] = bsrc[i];
  double2* l = (double2*)(smem + A.n_sb);
  if constexpr (LUTL)
    for (int i = threadIdx.x; i < A.lut_len; i += blockDim.x) l[i] = A.tlut[i];
  sb = smem;
  slut = l;
}

// k_hv_rows_fast with the terms in batches of B (fast_rows_acc).
template <typename I, typename VT, int M, int LK, int RPL, int B, int TW, bool LUTL>
__global__ void __launch_bounds__(256) k_hv_rows_fast_batch(HvArgs A) {
  // dynamic LDS: A.n_sb binomials (rows 0..N), then with LUTL the lut_len
  // coefficients (read per element: from LDS rather than through the vector
  // memory path)
  extern __shared__ int64_t fast_smem[];
  if (A.skip && *A.skip) return;
  const int64_t* sb;
  const double2* slut;
  fast_stage<LUTL>(A, fast_smem, sb, slut);
  __syncthreads();
  // Rows in runs per XCD: consecutive rows share their high bits, and so do
  // most of their targets and table entries, so each XCD's L2 sees fewer
  // distinct lines.  By default each XCD takes runs of 64 blocks (256 on
  // grids of 16,384 blocks or more) inside a window all XCDs share: measured
  // on the Heisenberg chain, L = 28 (1,412 blocks) 28.4 -> 28.1 us and L = 34
  // (67,108 blocks) 1.30 -> 1.155 ms against dispatch order.  xcd_rows (the
  // batch14x schedule) gives each XCD one run: fewest L2 misses, but 1.42 ms
  // at L = 34; the tuner times both.
  const int64_t bi = A.xcd_rows ? xcd_contiguous(blockIdx.x, gridDim.x)
                                : xcd_chunked(blockIdx.x, gridDim.x, gridDim.x >= 16384 ? 256 : 64);
  const int64_t r0 = bi * blockDim.x * RPL + threadIdx.x;
  if (r0 >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  int64_t rr[RPL];
  bool live[RPL];
  I s[RPL];
  int64_t ra[RPL];
  double rsq[RPL];
  double2 acc[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    rr[k] = r0 + (int64_t)k * blockDim.x;
    live[k] = rr[k] < A.Ns;
    const int64_t rc = live[k] ? rr[k] : r0;
    s[k] = ((const I*)A.basis)[rc];
    rsq[k] = 1.0;
    if constexpr (M != 0) rsq[k] = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, rc));
    ra[k] = comb_rank_sb<I>(sb, s[k]);
    acc[k] = make_double2(0.0, 0.0);
  }
  fast_rows_acc<I, VT, M, LK, RPL, B, TW, LUTL>(A, sb, slut, vin, s, ra, rsq, acc);
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    if (!live[k]) continue;
    const int64_t r = rr[k];
    double2 d = make_double2(0.0, 0.0);
    if (A.diag_arr) {
      if (A.diag_arr_cplx) d = ((const double2*)A.diag_arr)[r];
      else d.x = ((const double*)A.diag_arr)[r];
    }
    const double2 dsum = cadd(cmul(d, A.alpha_diag), A.shift);
    const double2 xr = VTr<VT>::ld(vin + r);
    double2 res;
    if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc[k]), cmul(dsum, xr));
    else res = make_double2(A.alpha_off.x * acc[k].x + dsum.x * xr.x, 0.0);
    if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
    VTr<VT>::st(vout + r, res);
  }
}

// One vector, G lanes per row: lane j of a row handles terms j, j+G, ... and
// the partial sums are combined with cross-lane shuffles (more waves and
// shorter per-lane dependency chains than one lane per row).
template <typename I, typename VT, int M, bool FERM, int LK, int G>
__global__ void __launch_bounds__(256) k_hv_rows_g(HvArgs A) {
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  if (A.skip && *A.skip) return;
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  constexpr bool CPLX = VTr<VT>::cplx;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = gid / G;
  const int lane = (int)(threadIdx.x % G);
  const bool live = r < A.Ns;
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  double2 acc = make_double2(0.0, 0.0);
  I s = 0;
  if (live) {
    s = ((const I*)A.basis)[r];
    double rsq = 1.0;
    if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
    const RowRank rr = row_rank_if<I, LK>(A, sb, s);
    for (int t = lane; t < A.n_off; t += G) {
      term_element<I, M, FERM, LK, CPLX>(A, sb, rr, A.off + t, s, rsq, A.off_conj, [&](int64_t c, double2 coef) {
        const double2 x = VTr<VT>::ld(vin + c);
        if constexpr (CPLX) acc = cadd(acc, cmul(coef, x));
        else acc.x += coef.x * x.x;
      });
    }
  }
  acc.x = grp_sum<G>(acc.x);
  if constexpr (CPLX) acc.y = grp_sum<G>(acc.y);
  if (!live || lane != 0) return;
  const double2 dsum = row_diag<I, M, FERM, LK, CPLX>(A, r, s);
  double2 res;
  if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc), cmul(dsum, VTr<VT>::ld(vin + r)));
  else res = make_double2(A.alpha_off.x * acc.x + dsum.x * VTr<VT>::ld(vin + r).x, 0.0);
  if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
  VTr<VT>::st(vout + r, res);
}

// One vector, rank-based lookups (table or no-symmetry rank), terms processed
// in batches of KB: all KB images and ranks first, then all KB table loads,
// then all KB (n, v) loads, so each lane keeps KB independent memory requests
// in flight instead of one dependent chain per term.
template <typename I, typename VT, int M, bool FERM, int LK, int KB>
__global__ void __launch_bounds__(256) k_hv_rows_batch(HvArgs A) {
  static_assert(LK == LOOKUP_TABLE || LK == LOOKUP_RANK, "rank-based lookups only");
  __shared__ int64_t sb[QSX_SB];
  if (A.skip && *A.skip) return;
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  constexpr bool CPLX = VTr<VT>::cplx;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.Ns) return;
  const I s = ((const I*)A.basis)[r];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, r));
  const RowRank rr = row_rank<I>(A.R, sb, s);
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const bool ferm = A.C->fermion != 0;
  double2 acc = make_double2(0.0, 0.0);
  for (int t0 = 0; t0 < A.n_off; t0 += KB) {
    int64_t rk[KB];
    double mg[KB];
    int ipw[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      rk[k] = -1;
      mg[k] = 0.0;
      ipw[k] = 0;
      const int t = t0 + k;
      if (t < A.n_off) {
        const TermDev* T = A.off + t;
        I s2 = s;
        double mag;
        int ip;
        if (apply_term<I>(T, ferm, s2, mag, ip)) {
          rk[k] = sector_rank_inc<I>(A.R, sb, s, rr, s2, T->lo, T->hi);
          mg[k] = mag;
          ipw[k] = ip;
        }
      }
    }
    int64_t c[KB];
    double2 fac[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      c[k] = -1;
      fac[k] = make_double2(1.0, 0.0);
      if constexpr (LK == LOOKUP_TABLE) {
        if (!table_entry<I>(A, rk[k], c[k], fac[k])) c[k] = -1;
      } else {
        c[k] = rk[k] < 0 ? (int64_t)-1 : A.R->Ns - 1 - rk[k];
      }
    }
    double2 x[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) x[k] = c[k] >= 0 ? VTr<VT>::ld(vin + c[k]) : make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      if (c[k] < 0) continue;
      const TermDev* T = A.off + t0 + k;
      double w = mg[k];
      if constexpr (M != 0) w *= rsq;
      if constexpr (CPLX) {
        double2 coef = cmul(make_double2(T->jr, T->ji), ipow_c(ipw[k]));
        coef = make_double2(coef.x * w, coef.y * w);
        if constexpr (M != 0) coef = cmul(coef, fac[k]);
        if (A.off_conj) coef = cconj(coef);
        acc = cadd(acc, cmul(coef, x[k]));
      } else {
        double v = T->jr * w * ((ipw[k] & 2) ? -1.0 : 1.0);
        if constexpr (M != 0) v *= fac[k].x;
        acc.x += v * x[k].x;
      }
    }
  }
  const double2 dsum = row_diag<I, M, FERM, LK, CPLX>(A, r, s);
  double2 res;
  if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc), cmul(dsum, VTr<VT>::ld(vin + r)));
  else res = make_double2(A.alpha_off.x * acc.x + dsum.x * VTr<VT>::ld(vin + r).x, 0.0);
  if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
  VTr<VT>::st(vout + r, res);
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
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  if (A.skip && *A.skip) return;
  stage_binom<LK>(A.R, sb);
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
      double2 fac;
      double2 val = make_double2(0.0, 0.0);
      if (resolve<I, M, FERM, LK>(A, sb, q_s[e], c, fac)) {
        const double2 x = VTr<VT>::ld(vin + c);
        if constexpr (CPLX) {
          double2 f = make_double2(1.0, 0.0);
          if constexpr (M != 0) f = fac;
          if (A.off_conj) f = cconj(f);
          val = cmul(f, x);
        } else {
          double f = 1.0;
          if constexpr (M != 0) f = fac.x;
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
  __shared__ int64_t sb[(LK == LOOKUP_TABLE || LK == LOOKUP_RANK) ? QSX_SB : 1];
  if (A.skip && *A.skip) return;
  stage_binom<LK>(A.R, sb);
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.Ns) return;
  constexpr bool CPLX = VTr<VT>::cplx;
  const I s = ((const I*)A.basis)[c];
  double rsq = 1.0;
  if constexpr (M != 0) rsq = 1.0 / sqrt(load_n(A.nrm, A.n_bytes, c));
  const RowRank rr = row_rank_if<I, LK>(A, sb, s);
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* vout = (VT*)A.v_out;
  const int64_t nv = A.nvec;
  for (int t = 0; t < A.n_off; ++t) {
    term_element<I, M, FERM, LK, CPLX>(A, sb, rr, A.off + t, s, rsq, A.off_conj, [&](int64_t r, double2 coef) {
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

// Fast-row launch.  RPL = 2 rows per lane for the direct rank (no table
// reads); with the table one row per lane keeps more waves in flight for the
// lookups.  nb terms per batch (1: k_hv_rows_fast; 4..8 or 14:
// k_hv_rows_fast_batch); nb = 0 picks the default measured on MI355X:
// batches (fast_batch_default) with the table, 1 for the direct rank, where
// the branch-free batches lose to skipping the gathers of terms that give a
// row nothing.
// Coefficient tables up to this many entries (16 KB) go to LDS.
constexpr int QSX_LUT_LDS = 1024;

template <typename I, typename VT, int M, int LK, int RPL, int TW, bool LUTL>
void launch_hv_fast_l(const HvArgs& A, int nb, size_t smem, hipStream_t st) {
  const int blk = 256;
  const dim3 grid((unsigned)qsx_grid(A.Ns, RPL * blk));
  switch (nb) {
    case 5: hipLaunchKernelGGL((k_hv_rows_fast_batch<I, VT, M, LK, RPL, 5, TW, LUTL>), grid, dim3(blk), smem, st, A); break;
    case 6: hipLaunchKernelGGL((k_hv_rows_fast_batch<I, VT, M, LK, RPL, 6, TW, LUTL>), grid, dim3(blk), smem, st, A); break;
    case 7: hipLaunchKernelGGL((k_hv_rows_fast_batch<I, VT, M, LK, RPL, 7, TW, LUTL>), grid, dim3(blk), smem, st, A); break;
    case 8: hipLaunchKernelGGL((k_hv_rows_fast_batch<I, VT, M, LK, RPL, 8, TW, LUTL>), grid, dim3(blk), smem, st, A); break;
    case 14: hipLaunchKernelGGL((k_hv_rows_fast_batch<I, VT, M, LK, RPL, 14, TW, LUTL>), grid, dim3(blk), smem, st, A); break;
    default: hipLaunchKernelGGL((k_hv_rows_fast_batch<I, VT, M, LK, RPL, 4, TW, LUTL>), grid, dim3(blk), smem, st, A); break;
  }
}

// Default batch for n terms: the fewest batches of at most 8 terms, then the
// smallest batch (4..8) that keeps that count, so the last batch is not
// padded more than needed (28 terms: 4 batches of 7).
inline int fast_batch_default(int n) {
  if (n <= 4) return 4;
  const int nbat = (n + 7) / 8;
  const int b = (n + nbat - 1) / nbat;
  return b < 4 ? 4 : b;
}

template <typename I, typename VT, int M, int LK, int RPL, int TW>
void launch_hv_fast_b(HvArgs A, int nb, hipStream_t st) {
  if (nb <= 1) {
    hipLaunchKernelGGL((k_hv_rows_fast<I, VT, M, LK, RPL>), dim3((unsigned)qsx_grid(A.Ns, RPL * 256)), dim3(256), 0,
                       st, A);
    return;
  }
  if (A.n_sb <= 0) A.n_sb = QSX_SB + 1;  // every row (the caller did not size it)
  const size_t sb_bytes = (size_t)A.n_sb * sizeof(int64_t);
  if constexpr (TW != 0) {
    if (A.tlut && A.lut_len > 0 && A.lut_len <= QSX_LUT_LDS) {
      launch_hv_fast_l<I, VT, M, LK, RPL, TW, true>(A, nb, sb_bytes + (size_t)A.lut_len * sizeof(double2), st);
      return;
    }
  }
  launch_hv_fast_l<I, VT, M, LK, RPL, TW, false>(A, nb, sb_bytes, st);
}

template <typename I, typename VT, int M, int LK>
void launch_hv_fast(const HvArgs& A, int nb, hipStream_t st) {
  if constexpr (LK == LOOKUP_RANK) {
    launch_hv_fast_b<I, VT, M, LK, 2, 0>(A, nb == 0 ? 1 : nb, st);
  } else {
    if (nb == 0) nb = fast_batch_default(A.n_off);
    if (A.table_u64) launch_hv_fast_b<I, VT, M, LK, 1, 8>(A, nb, st);
    else launch_hv_fast_b<I, VT, M, LK, 1, 4>(A, nb, st);
  }
}

// Dispatch over (maps, fermion sign, lookup, schedule, NV).
// schedule codes: 0 one lane per row, 1 compacted queue, 2/3/4 = 2/4/8 lanes
// per row, 5/6 = batches of 4/8 terms (rank-based lookups), 10 = batches of
// 14 terms on the fast-row path (elsewhere as 6), 12 = as 10 with each XCD's
// workgroups on one contiguous run of rows (A.xcd_rows)
template <typename I, typename VT, int M, bool FERM, int LK>
void launch_hv_one(const HvArgs& A, hipStream_t st) {
  const int blk = 256;
  const dim3 grid((unsigned)qsx_grid(A.Ns, blk));
  if (A.nvec == 1 && A.sched == 1 && A.n_off > 0)
    hipLaunchKernelGGL((k_hv_queue<I, VT, M, FERM, LK>), grid, dim3(QBLK), 0, st, A);
  else if (A.nvec == 1 && A.sched == 2)
    hipLaunchKernelGGL((k_hv_rows_g<I, VT, M, FERM, LK, 2>), dim3((unsigned)qsx_grid(A.Ns * 2, blk)), dim3(blk), 0, st, A);
  else if (A.nvec == 1 && A.sched == 3)
    hipLaunchKernelGGL((k_hv_rows_g<I, VT, M, FERM, LK, 4>), dim3((unsigned)qsx_grid(A.Ns * 4, blk)), dim3(blk), 0, st, A);
  else if (A.nvec == 1 && A.sched == 4)
    hipLaunchKernelGGL((k_hv_rows_g<I, VT, M, FERM, LK, 8>), dim3((unsigned)qsx_grid(A.Ns * 8, blk)), dim3(blk), 0, st, A);
  else if (A.nvec == 1 && A.fast_rows && (A.sched == 8 || A.sched == 9) && !FERM && LK == LOOKUP_RANK && M == 0) {
    if constexpr (!FERM && LK == LOOKUP_RANK && M == 0) {
      const dim3 g2((unsigned)qsx_grid(A.Ns, 2 * blk));
      if (A.sched == 8) hipLaunchKernelGGL((k_hv_rows_fast_chunk<I, VT, 2, 4>), g2, dim3(blk), 0, st, A);
      else hipLaunchKernelGGL((k_hv_rows_fast_chunk<I, VT, 2, 8>), g2, dim3(blk), 0, st, A);
    }
  }
  else if (A.nvec == 1 && A.fast_rows && (A.sched == 0 || A.sched == 5 || A.sched == 6 || A.sched == 10 || A.sched == 12) &&
           !FERM && (LK == LOOKUP_TABLE || LK == LOOKUP_RANK)) {
    if constexpr (!FERM && (LK == LOOKUP_TABLE || LK == LOOKUP_RANK))
      launch_hv_fast<I, VT, M, LK>(A, A.sched == 0 ? 0 : A.sched == 5 ? 4 : A.sched == 6 ? 8 : 14, st);
  }
  else if (A.nvec == 1 && (A.sched == 5 || A.sched == 6 || A.sched == 10 || A.sched == 12) &&
           (LK == LOOKUP_TABLE || LK == LOOKUP_RANK)) {
    if constexpr (LK == LOOKUP_TABLE || LK == LOOKUP_RANK) {
      if (A.sched == 5) hipLaunchKernelGGL((k_hv_rows_batch<I, VT, M, FERM, LK, 4>), grid, dim3(blk), 0, st, A);
      else hipLaunchKernelGGL((k_hv_rows_batch<I, VT, M, FERM, LK, 8>), grid, dim3(blk), 0, st, A);
    }
  }
  else if (A.nvec == 1)
    hipLaunchKernelGGL((k_hv_rows<I, VT, M, FERM, LK, 1>), grid, dim3(blk), 0, st, A);
  else
    hipLaunchKernelGGL((k_hv_rows<I, VT, M, FERM, LK, 0>), grid, dim3(blk), 0, st, A);
}

// Split-chain H.v (schedule 7; sector and terms as k_hv_rows_fast requires,
// no symmetry maps, N <= 32, one vector).  One workgroup per block u: the
// block's x is staged in LDS for the lower-bit terms, an upper-bit term
// reads the partner block's x at the same positions (consecutive lanes,
// consecutive addresses), the few terms across the halves gather from their
// target block.  Blocks are launched largest first and grouped by the
// popcount of u, so the partner blocks of the ones in flight belong to one
// popcount class (94 MB of complex64 at N = 28) and stay in the last-level
// cache.
constexpr int SPLIT_THREADS = 512;
constexpr int SPLIT_MAX_UPU = 64;
constexpr int SPLIT_MAX_K = 16;
constexpr int SPLIT_MAX_LUT = 1024;

template <typename VT>
__global__ void __launch_bounds__(SPLIT_THREADS) k_hv_split(HvArgs A) {
  extern __shared__ unsigned char split_smem[];
  __shared__ int64_t ub[SPLIT_MAX_UPU];
  __shared__ double2 uc[SPLIT_MAX_UPU];
  __shared__ int nu;
  __shared__ int64_t sbk[SPLIT_MAX_K * 33];  // C(n, k) for n < K (the ranks of d)
  __shared__ double2 lut[SPLIT_MAX_LUT];
  if (A.skip && *A.skip) return;
  const SplitDev S = *A.split;  // by value: the kernel's stores cannot alias it
  const uint32_t u = (uint32_t)S.order[blockIdx.x];
  const int K = S.K;
  const int Np = A.R->Np;
  const int c = Np - __builtin_popcount(u);
  const int off = S.cls_off[c];
  const int n = S.cls_off[c + 1] - off;
  const int64_t base = S.ublk[u];
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  VT* xs = (VT*)split_smem;
  constexpr bool CPLX = VTr<VT>::cplx;
  for (int p = threadIdx.x; p < n; p += blockDim.x) xs[p] = vin[base + p];
  for (int i = threadIdx.x; i < K * 33; i += blockDim.x) sbk[i] = (&A.R->binom[0][0])[i];
  for (int i = threadIdx.x; i < S.n_lut; i += blockDim.x) lut[i] = S.lut[i];
  // upper-bit terms that read upper bits only: the same partner block and
  // coefficient for every row of the block
  if (threadIdx.x == 0) {
    int k = 0;
    const uint32_t su = u << K;
    for (int t = 0; t < S.n_upu && k < SPLIT_MAX_UPU; ++t) {
      const TermDev* __restrict__ T = S.upu + t;
      const uint32_t v = su & (uint32_t)T->req_mask;
      if (v != (uint32_t)T->req_val && !(T->has2 && v == (uint32_t)T->req_val2)) continue;
      const uint32_t u2 = u ^ (uint32_t)(T->flip >> K);
      if (__builtin_popcount(u2) != Np - c) continue;
      const double w = (__builtin_popcount(su & (uint32_t)T->smask) & 1) ? -1.0 : 1.0;
      double2 coef = make_double2(T->cr * w, T->ci * w);
      if (A.off_conj) coef = cconj(coef);
      ub[k] = S.ublk[u2];
      uc[k] = coef;
      ++k;
    }
    nu = k;
  }
  __syncthreads();
  const uint32_t dmask = (1u << K) - 1u;
  const uint16_t* __restrict__ dl = S.dlist + off;
  const int nup = nu;
  const int ke = S.ell_k[c];
  const uint32_t* __restrict__ el = S.ell + S.ell_off[c];
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    const int64_t r = base + p;
    // the row's own reads first (diagonal), then the partner blocks' values
    // in batches of 8, all requested before any is used
    double2 dg = make_double2(0.0, 0.0);
    if (A.diag_arr) {
      if (A.diag_arr_cplx) dg = ((const double2*)A.diag_arr)[r];
      else dg.x = ((const double*)A.diag_arr)[r];
    }
    for (int j0 = 0; j0 < nup; j0 += 8) {
      double2 xv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        xv[q] = (j0 + q < nup) ? VTr<VT>::ld(vin + ub[j0 + q] + p) : make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (j0 + q >= nup) break;
        if constexpr (CPLX) acc = cadd(acc, cmul(uc[j0 + q], xv[q]));
        else acc.x += uc[j0 + q].x * xv[q].x;
      }
    }
    // lower half: the class's sparse matrix (entries read coalesced, 8 at a
    // time) against the staged block
    for (int k0 = 0; k0 < ke; k0 += 8) {
      uint32_t ev[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) ev[q] = (k0 + q < ke) ? el[(int64_t)(k0 + q) * n + p] : (uint32_t)p;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double2 x = VTr<VT>::ld(xs + (ev[q] & 0xFFFFu));
        const double2 cf = lut[ev[q] >> 16];
        if constexpr (CPLX) acc = cadd(acc, cmul(cf, x));
        else acc.x += cf.x * x.x;
      }
    }
    const uint32_t d = dl[p];
    const uint32_t s = (u << K) | d;
    const int64_t ra = (int64_t)n - 1 - p;  // ascending rank of d in its class
    auto add = [&](const TermDev* __restrict__ T, const double2 x) {
      const double w = (__builtin_popcount(s & (uint32_t)T->smask) & 1) ? -1.0 : 1.0;
      if constexpr (CPLX) {
        double2 coef = make_double2(T->cr * w, T->ci * w);
        if (A.off_conj) coef = cconj(coef);
        acc = cadd(acc, cmul(coef, x));
      } else {
        acc.x += T->cr * w * x.x;
      }
    };
    auto applies = [&](const TermDev* __restrict__ T) -> bool {
      const uint32_t v = s & (uint32_t)T->req_mask;
      return v == (uint32_t)T->req_val || (T->has2 && v == (uint32_t)T->req_val2);
    };
    // lower half: the partner position from the rank of d, the value from LDS
    for (int t = 0; t < S.n_low; ++t) {
      const TermDev* __restrict__ T = S.low + t;
      if (!applies(T)) continue;
      const uint32_t d2 = d ^ (uint32_t)T->flip;
      if (__builtin_popcount(d2) != c) continue;
      const int lo = T->lo, hi = T->hi;
      int64_t rk;
      if (hi == lo + 1) {
        const int64_t dl1 = sbk[lo * 33 + __builtin_popcount(d & ((1u << lo) - 1u))];
        rk = ((d >> lo) & 1u) ? ra + dl1 : ra - dl1;
      } else {
        rk = ra - window_rank<uint32_t>(sbk, d, lo, hi) + window_rank<uint32_t>(sbk, d2, lo, hi);
      }
      add(T, VTr<VT>::ld(xs + (n - 1 - rk)));
    }
    for (int t = 0; t < S.n_up; ++t) {
      const TermDev* __restrict__ T = S.up + t;
      if (!applies(T)) continue;
      const uint32_t u2 = u ^ (uint32_t)(T->flip >> K);
      if (__builtin_popcount(u2) != Np - c) continue;
      add(T, VTr<VT>::ld(vin + S.ublk[u2] + p));
    }
    for (int t = 0; t < S.n_cross; ++t) {
      const TermDev* __restrict__ T = S.cross + t;
      if (!applies(T)) continue;
      const uint32_t s2 = s ^ (uint32_t)T->flip;
      if (__builtin_popcount(s2) != Np) continue;
      const int64_t b2 = S.ublk[s2 >> K];
      if (b2 < 0) continue;
      add(T, VTr<VT>::ld(vin + b2 + S.dpos[s2 & dmask]));
    }
    const double2 dsum = cadd(cmul(dg, A.alpha_diag), A.shift);
    const double2 xr = VTr<VT>::ld(xs + p);
    double2 res;
    if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc), cmul(dsum, xr));
    else res = make_double2(A.alpha_off.x * acc.x + dsum.x * xr.x, 0.0);
    if (A.beta) res = cadd(res, VTr<VT>::ld(vout + r));
    VTr<VT>::st(vout + r, res);
  }
}

// Two-pass form of the split-chain layout (schedule 11).  Pass 1, one
// workgroup per two blocks u of one class (largest first): both blocks are
// staged in LDS (four loads per thread before their stores) and every row's
// lower-half terms are summed in FP64 (the class's sliced-ELL matrix, each
// entry serving both blocks, then the lower-flip terms that read upper bits)
// into Z, stored in the vector dtype.
//
// Two forms of the ELL loop.  KE > 0 (paired blocks, 512 threads): a row's
// entries are loaded in one round, the next row's while this one is summed,
// and the first row's before the staging barrier; a class wider than KE
// takes the KE = 0 loop.  KE = 0 (one block, 1,024 threads): entries in
// rounds of eight; about 60 VGPRs, so two workgroups of 16 waves fit a CU
// beside their LDS.  Measured on the XXZ L = 28 row pass (13 entries per
// row): complex64 pairs 311 -> 263-283 us with KE = 16, complex128 single
// blocks 448 -> 386 us with 1,024 threads.
constexpr int SPLIT2_ROWS_THREADS = 512;
constexpr int SPLIT2_ROWS_THREADS1 = 1024;  // one block per workgroup
constexpr int SPLIT2_ROWS_KE = 16;          // ELL entries per row loaded in one round (pairs)
constexpr int SPLIT2_UNIT_U = 16;  // u per column-strip unit (two or four u per wave)

// element coefficient of a bit-parallel term on state s (sign from smask)
template <bool CPLX>
__device__ __forceinline__ double2 split_coef(const HvArgs& A, const TermDev* __restrict__ T, uint32_t s) {
  const double w = (__builtin_popcount(s & (uint32_t)T->smask) & 1) ? -1.0 : 1.0;
  double2 coef = make_double2(T->cr * w, CPLX ? T->ci * w : 0.0);
  if (CPLX && A.off_conj) coef = cconj(coef);
  return coef;
}

__device__ __forceinline__ bool split_applies(const TermDev* __restrict__ T, uint32_t s) {
  const uint32_t v = s & (uint32_t)T->req_mask;
  return v == (uint32_t)T->req_val || (T->has2 && v == (uint32_t)T->req_val2);
}

template <typename VT, int R, int KE, int NT>
__global__ void __launch_bounds__(NT) k_split2_rows(HvArgs A) {
  extern __shared__ unsigned char split_smem[];
  __shared__ int64_t sbk[SPLIT_MAX_K * 33];  // C(n, k) for n < K (the ranks of d)
  __shared__ double2 lut[SPLIT_MAX_LUT];
  if (A.skip && *A.skip) return;
  const SplitDev S = *A.split;
  const int u1 = S.rpair[2 * blockIdx.x], u2 = R > 1 ? S.rpair[2 * blockIdx.x + 1] : -1;
  const int nb = u2 >= 0 ? 2 : 1;
  const uint32_t uq[2] = {(uint32_t)u1, (uint32_t)(u2 >= 0 ? u2 : u1)};
  const int K = S.K;
  const int Np = A.R->Np;
  const int c = Np - __builtin_popcount(uq[0]);
  const int n = S.cls_off[c + 1] - S.cls_off[c];
  const int64_t base[2] = {S.ublk[uq[0]], S.ublk[uq[1]]};
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ z = (VT*)S.z;
  VT* xs = (VT*)split_smem;  // block 0 at xs[0 .. n), block 1 at xs[n .. 2n)
  constexpr bool CPLX = VTr<VT>::cplx;
  const int ke = S.ell_k[c];
  const uint32_t* __restrict__ el = S.ell + S.ell_off[c];
  constexpr int KA = KE > 0 ? KE : 1;
  const bool pipe = KE > 0 && ke <= KE;
  uint32_t en[KA];  // the ELL entries of this thread's next row
  if (pipe) {
#pragma unroll
    for (int q = 0; q < KA; ++q)
      en[q] = (q < ke && (int)threadIdx.x < n) ? el[(int64_t)q * n + threadIdx.x] : 0u;
  }
  {
    // staging: four loads per thread before their LDS stores
    const int bd = blockDim.x, tot = nb * n;
    int i = threadIdx.x;
    for (; i + 3 * bd < tot; i += 4 * bd) {
      VT t[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = i + q * bd;
        t[q] = vin[(j < n ? base[0] : base[1] - n) + j];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) xs[i + q * bd] = t[q];
    }
    for (; i < tot; i += bd) xs[i] = vin[(i < n ? base[0] : base[1] - n) + i];
  }
  if (S.n_low > 0)
    for (int i = threadIdx.x; i < K * 33; i += blockDim.x) sbk[i] = (&A.R->binom[0][0])[i];
  for (int i = threadIdx.x; i < S.n_lut; i += blockDim.x) lut[i] = S.lut[i];
  __syncthreads();
  const uint16_t* __restrict__ dl = S.dlist + S.cls_off[c];
  const int xo = nb > 1 ? n : 0;  // second block's offset in LDS (the first again when alone)
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    double2 acc0 = make_double2(0.0, 0.0), acc1 = make_double2(0.0, 0.0);
    if (pipe) {
      uint32_t ev[KA];
      const int pn = p + (int)blockDim.x;
#pragma unroll
      for (int q = 0; q < KA; ++q) {
        ev[q] = en[q];
        en[q] = (q < ke && pn < n) ? el[(int64_t)q * n + pn] : 0u;
      }
#pragma unroll
      for (int q = 0; q < KA; ++q) {
        if (q >= ke) break;
        const int pos = (int)(ev[q] & 0xFFFFu);
        const double2 cf = lut[ev[q] >> 16];
        const double2 x0 = VTr<VT>::ld(xs + pos);
        if constexpr (CPLX) acc0 = cadd(acc0, cmul(cf, x0));
        else acc0.x += cf.x * x0.x;
        if constexpr (R > 1) {
          const double2 x1 = VTr<VT>::ld(xs + xo + pos);
          if constexpr (CPLX) acc1 = cadd(acc1, cmul(cf, x1));
          else acc1.x += cf.x * x1.x;
        }
      }
    } else
    for (int k0 = 0; k0 < ke; k0 += 8) {
      uint32_t ev[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) ev[q] = (k0 + q < ke) ? el[(int64_t)(k0 + q) * n + p] : (uint32_t)p;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int pos = (int)(ev[q] & 0xFFFFu);
        const double2 cf = lut[ev[q] >> 16];
        const double2 x0 = VTr<VT>::ld(xs + pos);
        if constexpr (CPLX) acc0 = cadd(acc0, cmul(cf, x0));
        else acc0.x += cf.x * x0.x;
        if constexpr (R > 1) {
          const double2 x1 = VTr<VT>::ld(xs + xo + pos);
          if constexpr (CPLX) acc1 = cadd(acc1, cmul(cf, x1));
          else acc1.x += cf.x * x1.x;
        }
      }
    }
    if (S.n_low > 0) {
      const uint32_t d = dl[p];
      const int64_t ra = (int64_t)n - 1 - p;  // ascending rank of d in its class
      for (int b = 0; b < nb; ++b) {
        const uint32_t s = (uq[b] << K) | d;
        double2 acc = b ? acc1 : acc0;
        for (int t = 0; t < S.n_low; ++t) {
          const TermDev* __restrict__ T = S.low + t;
          if (!split_applies(T, s)) continue;
          const uint32_t d2 = d ^ (uint32_t)T->flip;
          if (__builtin_popcount(d2) != c) continue;
          const int lo = T->lo, hi = T->hi;
          int64_t rk;
          if (hi == lo + 1) {
            const int64_t dl1 = sbk[lo * 33 + __builtin_popcount(d & ((1u << lo) - 1u))];
            rk = ((d >> lo) & 1u) ? ra + dl1 : ra - dl1;
          } else {
            rk = ra - window_rank<uint32_t>(sbk, d, lo, hi) + window_rank<uint32_t>(sbk, d2, lo, hi);
          }
          const double2 x = VTr<VT>::ld(xs + (b ? xo : 0) + (n - 1 - rk));
          const double2 cf = split_coef<CPLX>(A, T, s);
          if constexpr (CPLX) acc = cadd(acc, cmul(cf, x));
          else acc.x += cf.x * x.x;
        }
        if (b) acc1 = acc; else acc0 = acc;
      }
    }
    st_stream<VT>(z + base[0] + p, acc0);
    if (nb > 1) st_stream<VT>(z + base[1] + p, acc1);
  }
}

// Pass 2, the column-strip pass: unit = (class c, strip of 64 positions, 16
// u of the class), strip-major and each XCD walking a contiguous run of
// units, so the strip of every u of the class (C(K, c) u x 64 elements) stays
// in that XCD's L2 while the upper-half terms read their partner blocks at
// the same positions.  Wave w takes RW u (wave-uniform: their partner
// blocks and coefficients are scalar), lane l position 64 s + l.  Then the
// upper-flip terms that read lower bits and the crossing terms (per lane),
// the diagonal and the output: y = alpha_off (Z + those sums) + (alpha_diag
// diag + shift) x (+ y).
__device__ __forceinline__ int64_t rl_i64(int64_t x, int t) {
  const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), t);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), t);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double rl_dbl(double x, int t) { return __longlong_as_double(rl_i64(__double_as_longlong(x), t)); }

template <typename VT, int HD, int XB, bool EPI, int RW, int WPE = 1>
__global__ void __launch_bounds__(64 * SPLIT2_UNIT_U / RW) __attribute__((amdgpu_waves_per_eu(WPE)))
k_split2_strips(HvArgs A) {
  if (A.skip && *A.skip) return;
  const SplitDev S = *A.split;
  constexpr int NW = SPLIT2_UNIT_U / RW;  // waves per workgroup
  constexpr bool CPLX = VTr<VT>::cplx;
  const int K = S.K, Np = A.R->Np;
  const VT* __restrict__ vin = (const VT*)A.v_in;
  VT* __restrict__ vout = (VT*)A.v_out;
  const VT* __restrict__ z = (const VT*)S.z;
  // class tables and the up / crossing terms in LDS (one round per workgroup)
  __shared__ int64_t s_uoff[SPLIT_MAX_K + 2];
  __shared__ int32_t s_cls[SPLIT_MAX_K + 2], s_ocl[SPLIT_MAX_K + 2];
  __shared__ SplitXT s_xt[SPLIT2_MAX_XT];
  const int nxt = S.n_up + S.n_cross;
  if (threadIdx.x < K + 2) {
    s_uoff[threadIdx.x] = S.uoff[threadIdx.x];
    s_cls[threadIdx.x] = S.cls_off[threadIdx.x];
    s_ocl[threadIdx.x] = S.ocls[threadIdx.x];
  }
  if (threadIdx.x < nxt) s_xt[threadIdx.x] = S.xt[threadIdx.x];
  __syncthreads();
  // the wave index as a scalar, so the rows' bases and coefficients stay scalar
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const uint32_t dmask = (1u << K) - 1u;
  const int upw = S.upw;
  double maxB = 0.0, maxF = 0.0;  // EPI: this thread's largest |term| and |F|
  for (int64_t unit = xcd_contiguous(blockIdx.x, gridDim.x); unit < S.nunits; unit += gridDim.x) {
    int c = 0;
    while (c < K && s_uoff[c + 1] <= unit) ++c;
    const int n = s_cls[c + 1] - s_cls[c];
    const int cnt = s_ocl[c + 1] - s_ocl[c];
    const int64_t ngrp = (cnt + SPLIT2_UNIT_U - 1) / SPLIT2_UNIT_U;
    const int64_t loc = unit - s_uoff[c];
    const int64_t sp = loc / ngrp, g = loc - sp * ngrp;
    const int i0 = (int)(g * SPLIT2_UNIT_U) + w * RW;
    if (i0 >= cnt) continue;
    const int nr = cnt - i0 < RW ? cnt - i0 : RW;
    const int p = (int)(sp * 64) + lane;
    const bool pok = p < n;
    const int ps = pok ? p : n - 1;  // lanes past the class read in bounds, store nothing
    // one round: the position's d, the rows' u, first rows and partner lists
    // (lane t holds entry t of each row's list)
    const uint32_t d = S.dlist[s_cls[c] + ps];
    uint32_t uu[RW];
    int64_t bs[RW], pb[RW], xbl[RW];
    int npart[RW];
    double2 pc[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int64_t ii = (int64_t)s_ocl[c] + i0 + (r < nr ? r : 0);
      uu[r] = (uint32_t)S.ord2[ii];
      bs[r] = S.obase[ii];
      npart[r] = r < nr ? S.upn[ii] : 0;
      pb[r] = lane < upw ? S.upb[ii * upw + lane] : -1;
      pc[r] = lane < upw ? S.upc[ii * upw + lane] : make_double2(0.0, 0.0);
      xbl[r] = lane < nxt ? S.xb[ii * nxt + lane] : -1;
    }
    double2 acc[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) acc[r] = ld_stream<VT>(z + bs[r] + ps);
    int nmax = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) nmax = npart[r] > nmax ? npart[r] : nmax;
    // upper-half terms reading upper bits only: partner block, same position,
    // eight loads in flight
    for (int t0 = 0; t0 < nmax; t0 += HD) {
      VT v[HD][RW];
#pragma unroll
      for (int h = 0; h < HD; ++h) {
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const int t = t0 + h;
          const bool on = t < npart[r];  // wave-uniform
          const int64_t b2 = on ? rl_i64(pb[r], t) : bs[r];
          v[h][r] = vin[b2 + ps];
        }
      }
#pragma unroll
      for (int h = 0; h < HD; ++h)
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const int t = t0 + h;
          if (t >= npart[r]) continue;  // wave-uniform
          const double2 cf = make_double2(rl_dbl(pc[r].x, t), rl_dbl(pc[r].y, t));
          const double2 x = VTr<VT>::ld(&v[h][r]);
          if constexpr (CPLX) acc[r] = cadd(acc[r], cmul(cf, x));
          else acc[r].x += cf.x * x.x;
        }
    }
    // upper-flip terms reading lower bits (partner position = own position)
    // and crossing terms (partner position from dpos): per-lane conditions;
    // for each term the rows' position reads, then their partner reads
    for (int t0 = 0; t0 < nxt; t0 += XB) {
      int64_t a[XB][RW];
      bool on[XB][RW];
#pragma unroll
      for (int h = 0; h < XB; ++h) {
        const int t = t0 + h < nxt ? t0 + h : nxt - 1;
        const SplitXT T = s_xt[t];
        const bool upt = t < S.n_up;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const uint32_t s = (uu[r] << K) | d;
          const uint32_t s2 = s ^ T.flip;
          const uint32_t vv = s & T.rm;
          const int64_t b2 = rl_i64(xbl[r], t);  // wave-uniform
          on[h][r] = t0 + h < nxt && r < nr && pok && b2 >= 0 && (vv == T.rv || (T.has2 && vv == T.rv2)) &&
                     __builtin_popcount(s2) == Np;
          a[h][r] = upt ? (int64_t)ps : (int64_t)S.dpos[s2 & dmask];
          a[h][r] += on[h][r] ? b2 : bs[r];
          if (!on[h][r]) a[h][r] = bs[r] + ps;
        }
      }
      VT v[XB][RW];
#pragma unroll
      for (int h = 0; h < XB; ++h)
#pragma unroll
        for (int r = 0; r < RW; ++r) v[h][r] = vin[a[h][r]];
#pragma unroll
      for (int h = 0; h < XB; ++h) {
        const int t = t0 + h < nxt ? t0 + h : nxt - 1;
        const SplitXT T = s_xt[t];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const uint32_t s = (uu[r] << K) | d;
          const double wt = on[h][r] ? ((__builtin_popcount(s & T.sm) & 1) ? -1.0 : 1.0) : 0.0;
          double2 cf = make_double2(T.cr * wt, CPLX ? T.ci * wt : 0.0);
          if (CPLX && A.off_conj) cf = cconj(cf);
          const double2 x = VTr<VT>::ld(&v[h][r]);
          if constexpr (CPLX) acc[r] = cadd(acc[r], cmul(cf, x));
          else acc[r].x += cf.x * x.x;
        }
      }
    }
    if (!pok) continue;
    double2 xr[RW], dg[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int64_t row = bs[r] + p;
      xr[r] = VTr<VT>::ld(vin + row);
      dg[r] = make_double2(0.0, 0.0);
      if (A.diag_arr) {
        if (A.diag_arr_cplx) dg[r] = ld_stream<double2>((const double2*)A.diag_arr + row);
        else dg[r] = ld_stream<double>((const double*)A.diag_arr + row);
      }
    }
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      if (r >= nr) break;
      const int64_t row = bs[r] + p;
      const double2 dsum = cadd(cmul(dg[r], A.alpha_diag), A.shift);
      double2 res;
      if constexpr (CPLX) res = cadd(cmul(A.alpha_off, acc[r]), cmul(dsum, xr[r]));
      else res = make_double2(A.alpha_off.x * acc[r].x + dsum.x * xr[r].x, 0.0);
      if (A.beta) res = cadd(res, VTr<VT>::ld(vout + row));
      st_stream<VT>(vout + row, res);
      if constexpr (EPI) {
        // the stored (rounded) term enters F, as the separate epilogue does
        VT* __restrict__ F = (VT*)A.epi_F;
        const double2 rs = VTr<VT>::rnd(res);
        const double2 f = VTr<VT>::rnd(cadd(ld_stream<VT>(F + row), rs));
        st_stream<VT>(F + row, f);
        maxB = fmax(maxB, CPLX ? sqrt(rs.x * rs.x + rs.y * rs.y) : fabs(rs.x));
        maxF = fmax(maxF, CPLX ? sqrt(f.x * f.x + f.y * f.y) : fabs(f.x));
      }
    }
  }
  if constexpr (EPI) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      maxB = fmax(maxB, __shfl_xor(maxB, o, 64));
      maxF = fmax(maxF, __shfl_xor(maxF, o, 64));
    }
    __shared__ double s_m[2][NW];
    if (lane == 0) {
      s_m[0][w] = maxB;
      s_m[1][w] = maxF;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double mb = s_m[0][0], mf = s_m[1][0];
      for (int i = 1; i < NW; ++i) {
        mb = fmax(mb, s_m[0][i]);
        mf = fmax(mf, s_m[1][i]);
      }
      // non-negative doubles order as their bit patterns
      atomicMax((unsigned long long*)A.epi_nB, (unsigned long long)__double_as_longlong(mb));
      atomicMax((unsigned long long*)A.epi_nF, (unsigned long long)__double_as_longlong(mf));
    }
  }
}

template <typename VT>
int launch_hv_split2(const HvArgs& A, hipStream_t st) {
  if (A.split_pairs)
    hipLaunchKernelGGL((k_split2_rows<VT, 2, SPLIT2_ROWS_KE, SPLIT2_ROWS_THREADS>), dim3((unsigned)A.split_nrp),
                       dim3(SPLIT2_ROWS_THREADS), (size_t)A.split_lds2, st, A);
  else
    hipLaunchKernelGGL((k_split2_rows<VT, 1, 0, SPLIT2_ROWS_THREADS1>), dim3((unsigned)A.split_nrp),
                       dim3(SPLIT2_ROWS_THREADS1), (size_t)A.split_lds2, st, A);
  const int64_t nb = A.split_nunits < 65536 ? (A.split_nunits > 0 ? A.split_nunits : 1) : 65536;
  // Two u per wave and eight waves (68-76 VGPRs, against 105-127 with four
  // u per wave and four waves), four upper-half terms per round of loads for
  // elements of at most 8 bytes and two for complex128, two crossing terms
  // per round.  Measured on the XXZ L = 28 strip pass: complex64 716.7 ->
  // 679.9 us (fused epilogue 911.4 -> 841.2 us), complex128 912.1 -> 892.6
  // us (four upper-half terms: 1,006 us).  The complex128 fused epilogue
  // needs 81 VGPRs with two u per wave (two workgroups per CU: 1,293 us,
  // against 1,179 us with four u per wave); held to 80 VGPRs (six waves per
  // SIMD, three workgroups) it takes 1,134 us.  Eight waves per SIMD spill.
  constexpr int HD = sizeof(VT) <= 8 ? 4 : 2;
  if (A.epi_F) {
    if constexpr (sizeof(VT) > 8)
      hipLaunchKernelGGL((k_split2_strips<VT, 2, 2, true, 2, 6>), dim3((unsigned)nb), dim3(64 * SPLIT2_UNIT_U / 2), 0,
                         st, A);
    else
      hipLaunchKernelGGL((k_split2_strips<VT, HD, 2, true, 2>), dim3((unsigned)nb), dim3(64 * SPLIT2_UNIT_U / 2), 0,
                         st, A);
  } else {
    hipLaunchKernelGGL((k_split2_strips<VT, HD, 2, false, 2>), dim3((unsigned)nb), dim3(64 * SPLIT2_UNIT_U / 2), 0,
                       st, A);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    qsx_set_error(std::string("split-chain launch failed: ") + hipGetErrorString(e));
    return QSX_ERR_HIP;
  }
  return QSX_OK;
}

template <typename VT>
int launch_hv_split(const HvArgs& A, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_hv_split<VT>, dim3((unsigned)nblk), dim3(SPLIT_THREADS), (size_t)A.split_lds, st, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    qsx_set_error(std::string("split-chain launch failed: ") + hipGetErrorString(e));
    return QSX_ERR_HIP;
  }
  return QSX_OK;
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
    if (!scatter && A.split && A.sched == 7 && A.fast_rows && A.nvec == 1)                     \
      return launch_hv_split<VT>(A, A.split_nblk, st);                                        \
    if (!scatter && A.split && A.sched == 11 && A.fast_rows && A.nvec == 1)                    \
      return launch_hv_split2<VT>(A, st);                                                     \
    if (A.epi_F) {                                                                             \
      qsx_set_error("the fused expm epilogue needs the two-pass split schedule");              \
      return QSX_ERR_UNSUPPORTED;                                                              \
    }                                                                                          \
    if (scatter) {                                                                             \
      if (N <= 32) return launch_hv_scatter<uint32_t, VT>(A, nmaps, fm, lookup, st);           \
      return launch_hv_scatter<uint64_t, VT>(A, nmaps, fm, lookup, st);                        \
    }                                                                                          \
    if (N <= 32) return launch_hv_rows<uint32_t, VT>(A, nmaps, fm, lookup, st);                \
    return launch_hv_rows<uint64_t, VT>(A, nmaps, fm, lookup, st);                             \
  }
