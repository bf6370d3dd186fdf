// Internal device/host structures of libquspin_hip (not part of the C-ABI).
//
// Bit convention: site j of an N-site lattice is bit N-1-j of the state
// integer, 1 = up/occupied (sphinx/source/tutorials/user_basis.rst:56-79).
// Symmetry maps act by pull, (g.s)[i] = s[map[i]] (xor inversion), and are
// compiled on the host into: optional full-word bit reverse, then a short
// list of (mask, shift) groups (all bits of a group move by the same
// distance), then an XOR with the inversion mask.  Translations on chains and
// ladders compile to 2 groups, parity to a reverse plus 1 group, spin
// inversion to 1 group plus an XOR.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QSX_MAX_MAPS 32
#define QSX_MAXN 64
#define QSX_MAX_OPS 16
#define QSX_MAX_TERMS 4096

enum { OP_I = 0, OP_Z = 1, OP_N = 2, OP_P = 3, OP_M = 4, OP_X = 5, OP_Y = 6 };
enum { KIND_HCB = 0, KIND_SPINLESS = 1, KIND_SPINFUL = 2 };
enum { LOOKUP_SEARCH = 0, LOOKUP_RANK = 1, LOOKUP_TABLE = 2 };

struct MapDev {
  uint64_t inv;   // XOR applied after the permutation
  int32_t rev;    // 1: bit-reverse the N-bit window first
  int32_t ng;     // number of (mask, shift) groups
  int32_t per;    // period of the map
  int32_t goff;   // first group in CoreDev::gmask / gshl / gshr
};

struct CoreDev {
  int32_t N;          // number of bits (2L for spinful fermions)
  int32_t kind;       // KIND_*
  int32_t nmaps;
  int32_t G;          // product of periods
  int32_t L;          // sites per species for spinful fermions, else N
  int32_t double_occ;
  int32_t fermion;    // 1 for fermion kinds
  int32_t pad;
  MapDev maps[QSX_MAX_MAPS];
  uint64_t gmask[QSX_MAX_MAPS * QSX_MAXN];
  int32_t gshl[QSX_MAX_MAPS * QSX_MAXN];
  int32_t gshr[QSX_MAX_MAPS * QSX_MAXN];
  // fermion re-ordering masks: for map k and source bit b,
  // wmask[k*64+b] = bits b' whose relative order with b flips under map k.
  uint64_t wmask[QSX_MAX_MAPS * QSX_MAXN];
};

// One operator string, ops in APPLICATION order (rightmost character first).
struct TermDev {
  double jr, ji;       // coupling
  uint64_t flip;       // net flip mask
  int32_t nops;
  int16_t lo, hi;      // lowest / highest flipped bit (window of the incremental rank)
  uint8_t code[QSX_MAX_OPS];
  uint8_t pos[QSX_MAX_OPS];  // bit positions
  // bit-parallel form for spin/hard-core-boson terms acting on distinct
  // sites: non-zero iff (s & req_mask) == req_val (or == req_val2 when has2:
  // two merged terms that differ only in the bits they require, e.g.
  // S+_i S-_j and S-_i S+_j with one coupling); amplitude
  // mag0 * (-1)^(nz - popc(s & zmask)) * i^(3 ny - 2 popc(s & ymask))
  uint64_t req_mask, req_val, zmask, ymask, req_val2;
  double mag0;
  int32_t fast, nz, ny, has2;
  // the same element with the state-independent factors folded:
  // J * amplitude = (cr + i ci) * (-1)^popc(s & smask), where
  // cr + i ci = J mag0 (-1)^nz i^(3 ny) and smask = zmask ^ ymask
  // (i^(-2 p) = (-1)^p, and popc(a) + popc(b) = popc(a ^ b) mod 2)
  double cr, ci;
  uint64_t smask;
};

// Sector description used by LOOKUP_RANK (combinadic index, no symmetry).
struct RankDev {
  int32_t mode;      // 0: hcb single sector, 1: spinful single (Nup,Ndown), 2: all 2^N states
  int32_t N, L;
  int32_t Np, Nup, Ndn;
  int64_t Ns;
  int64_t cdn;       // C(L, Ndn) for spinful
  int64_t binom[65][33];  // binom[n][k] for k <= 32 (sector popcounts <= 32)
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
template <typename I> __device__ __forceinline__ I bitrev_n(I x, int N);
template <> __device__ __forceinline__ uint32_t bitrev_n<uint32_t>(uint32_t x, int N) {
  return __builtin_bitreverse32(x) >> (32 - N);
}
template <> __device__ __forceinline__ uint64_t bitrev_n<uint64_t>(uint64_t x, int N) {
  return __builtin_bitreverse64(x) >> (64 - N);
}

template <typename I> __device__ __forceinline__ int popc(I x);
template <> __device__ __forceinline__ int popc<uint32_t>(uint32_t x) { return __builtin_popcount(x); }
template <> __device__ __forceinline__ int popc<uint64_t>(uint64_t x) { return __builtin_popcountll(x); }

template <typename I>
__device__ __forceinline__ I apply_map(const CoreDev* __restrict__ C, int k, I s) {
  const MapDev& m = C->maps[k];
  I x = m.rev ? bitrev_n<I>(s, C->N) : s;
  I y = 0;
  const int g0 = m.goff, ng = m.ng;
  for (int g = 0; g < ng; ++g) {
    const I msk = (I)C->gmask[g0 + g];
    y |= ((x & msk) << C->gshl[g0 + g]) >> C->gshr[g0 + g];
  }
  return y ^ (I)m.inv;
}

// (-1)^(number of fermion pairs whose order map k swaps) for state s.
template <typename I>
__device__ __forceinline__ int map_sign(const CoreDev* __restrict__ C, int k, I s) {
  int par = 0;
  I t = s;
  while (t) {
    int b = (sizeof(I) == 8) ? __builtin_ctzll((uint64_t)t) : __builtin_ctz((uint32_t)t);
    t &= t - 1;
    par += popc<I>(s & (I)C->wmask[k * QSX_MAXN + b]);
  }
  return (par & 1) ? -1 : 1;
}

// Orbit traversal over the group generated by the maps.  Elements are
// visited with map 0 outermost; the visitor gets (image, mixed-radix code with
// map 0 most significant, fermion sign) and returns false to stop early.
template <typename I, int K, int M, bool FERM, typename V>
__device__ __forceinline__ bool visit_rec(const CoreDev* __restrict__ C, I t, int code, int sign,
                                          V& vis) {
  if constexpr (K == M) {
    return vis(t, code, sign);
  } else {
    const int per = C->maps[K].per;
    for (int l = 0; l < per; ++l) {
      if (!visit_rec<I, K + 1, M, FERM, V>(C, t, code * per + l, sign, vis)) return false;
      if (l + 1 < per) {
        if constexpr (FERM) sign *= map_sign<I>(C, K, t);
        t = apply_map<I>(C, K, t);
      }
    }
    return true;
  }
}

// Any number of maps (odometer over small local arrays).
template <typename I, bool FERM, typename V>
__device__ bool visit_generic(const CoreDev* __restrict__ C, I s, V& vis) {
  const int M = C->nmaps;
  I t[QSX_MAX_MAPS + 1];
  int sg[QSX_MAX_MAPS + 1];
  int l[QSX_MAX_MAPS];
  for (int k = 0; k <= M; ++k) { t[k] = s; sg[k] = 1; }
  for (int k = 0; k < M; ++k) l[k] = 0;
  while (true) {
    int code = 0;
    for (int k = 0; k < M; ++k) code = code * C->maps[k].per + l[k];
    if (!vis(t[M], code, sg[M])) return false;
    int k = M - 1;
    for (; k >= 0; --k) {
      if (l[k] + 1 < C->maps[k].per) {
        if (FERM) sg[k + 1] *= map_sign<I>(C, k, t[k + 1]);
        t[k + 1] = apply_map<I>(C, k, t[k + 1]);
        l[k] += 1;
        for (int j = k + 1; j < M; ++j) { l[j] = 0; t[j + 1] = t[j]; sg[j + 1] = sg[j]; }
        break;
      }
    }
    if (k < 0) return true;
  }
}

template <typename I, int M, bool FERM, typename V>
__device__ __forceinline__ bool visit_orbit(const CoreDev* __restrict__ C, I s, V& vis) {
  if constexpr (M < 0) return visit_generic<I, FERM, V>(C, s, vis);
  else return visit_rec<I, 0, M, FERM, V>(C, s, 0, 1, vis);
}

// Representative search: orbit maximum, code of the first element reaching it,
// number of elements reaching it (= stabiliser size of the representative)
// and the fermion sign of that element.
struct OrbitRes {
  int code;
  int cnt;
  int sign;
};

template <typename I, int M, bool FERM>
__device__ __forceinline__ I orbit_max(const CoreDev* __restrict__ C, I s, OrbitRes& R) {
  I best = s;
  R.code = 0; R.cnt = 0; R.sign = 1;
  auto vis = [&](I t, int code, int sign) -> bool {
    if (t > best) { best = t; R.code = code; R.cnt = 1; R.sign = sign; }
    else if (t == best) { R.cnt += 1; }
    return true;
  };
  visit_orbit<I, M, FERM>(C, s, vis);
  return best;
}

// Apply one operator string to state s (in place).  Returns false when the
// amplitude vanishes.  Amplitude = mag * i^ipow.
template <typename I>
__device__ __forceinline__ bool apply_term(const TermDev* __restrict__ T, bool fermion, I& s,
                                           double& mag, int& ipow) {
  if (T->fast) {
    const I v = s & (I)T->req_mask;
    if (v != (I)T->req_val && !(T->has2 && v == (I)T->req_val2)) return false;
    const int zs = T->nz - popc<I>(s & (I)T->zmask);
    mag = (zs & 1) ? -T->mag0 : T->mag0;
    ipow = 3 * T->ny - 2 * popc<I>(s & (I)T->ymask);
    s ^= (I)T->flip;
    return true;
  }
  mag = 1.0;
  ipow = 0;
  const int n = T->nops;
  for (int k = 0; k < n; ++k) {
    const int p = T->pos[k];
    const int c = T->code[k];
    const bool b = (s >> p) & 1;
    switch (c) {
      case OP_I: break;
      case OP_Z: mag *= b ? 0.5 : -0.5; break;
      case OP_N: if (!b) return false; break;
      case OP_P:
        if (b) return false;
        if (fermion && (popc<I>(s >> (p + 1)) & 1)) mag = -mag;
        s ^= ((I)1 << p);
        break;
      case OP_M:
        if (!b) return false;
        if (fermion && (popc<I>(s >> (p + 1)) & 1)) mag = -mag;
        s ^= ((I)1 << p);
        break;
      case OP_X:
        if (fermion) { if (popc<I>(s >> (p + 1)) & 1) mag = -mag; }
        else mag *= 0.5;
        s ^= ((I)1 << p);
        break;
      case OP_Y:
        if (fermion) {
          // -i(c - c^dag): bit 0 -> +i, bit 1 -> -i
          if (popc<I>(s >> (p + 1)) & 1) mag = -mag;
          ipow += b ? 3 : 1;
        } else {
          // S^y: bit 0 -> -i/2, bit 1 -> +i/2
          mag *= 0.5;
          ipow += b ? 1 : 3;
        }
        s ^= ((I)1 << p);
        break;
      default: return false;
    }
  }
  return true;
}

// Binary search in a DESCENDING array on [lo, hi).  Returns index or -1.
template <typename I>
__device__ __forceinline__ int64_t bsearch_desc(const I* __restrict__ b, int64_t lo, int64_t hi, I key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const I v = b[mid];
    if (v == key) return mid;
    if (v > key) lo = mid + 1; else hi = mid;
  }
  return -1;
}

// Combinadic rank of an N-bit word with k set bits, ASCENDING order among
// words of that popcount (sum over set bits, j-th lowest at position p:
// C(p, j)).
template <typename I>
__device__ __forceinline__ int64_t comb_rank(const RankDev* __restrict__ R, I s) {
  int64_t r = 0;
  int j = 1;
  while (s) {
    const int p = (sizeof(I) == 8) ? __builtin_ctzll((uint64_t)s) : __builtin_ctz((uint32_t)s);
    s &= s - 1;
    r += R->binom[p][j];
    ++j;
  }
  return r;
}
