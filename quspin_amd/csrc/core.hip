// Basis core: symmetry-map compilation, device basis enumeration
// (make_basis), prefix blocks, representative and normalization.
//
// Replaces the native ``*_basis_core_wrap`` constructors and their
// make_basis / make_basis_blocks / representative / normalization methods
// (call sites: src/quspin/basis/basis_general/base_hcb.py:74-76,
// base_general.py:864-867, :945-947, :1085-1117, :1158, :1178;
// fermion.py:169-171, :543-545).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "qsx_host.h"

static thread_local std::string g_last_error;
void qsx_set_error(const std::string& msg) { g_last_error = msg; }

extern "C" const char* qsx_last_error(void) { return g_last_error.c_str(); }

extern "C" int qsx_version(void) { return 1; }

// ---------------------------------------------------------------------------
// host: map compilation
// ---------------------------------------------------------------------------
namespace {

int64_t gcd64(int64_t a, int64_t b) { while (b) { int64_t t = a % b; a = b; b = t; } return a < 0 ? -a : a; }

// Compile the pull map of one generator into (rev, groups, inv).
bool compile_map(const int32_t* map, int N, CoreDev& C, int k, int& goff) {
  std::vector<int> perm(N), inv(N);
  for (int i = 0; i < N; ++i) {
    int m = map[i];
    if (m < 0) { perm[i] = -(m + 1); inv[i] = 1; } else { perm[i] = m; inv[i] = 0; }
    if (perm[i] < 0 || perm[i] >= N) return false;
  }
  std::vector<int> seen(N, 0);
  for (int i = 0; i < N; ++i) { if (seen[perm[i]]) return false; seen[perm[i]] = 1; }
  uint64_t invmask = 0;
  // dest bit d = N-1-i pulls src bit N-1-perm[i]
  auto groups_for = [&](bool rev, std::vector<std::pair<int, uint64_t>>& out) {
    out.clear();
    for (int i = 0; i < N; ++i) {
      int d = N - 1 - i;
      int src = N - 1 - perm[i];
      if (rev) src = N - 1 - src;
      int delta = d - src;
      auto it = std::find_if(out.begin(), out.end(), [&](const std::pair<int, uint64_t>& g) { return g.first == delta; });
      if (it == out.end()) out.push_back({delta, 1ull << src});
      else it->second |= 1ull << src;
    }
  };
  for (int i = 0; i < N; ++i) if (inv[i]) invmask |= 1ull << (N - 1 - i);
  std::vector<std::pair<int, uint64_t>> gd, gr;
  groups_for(false, gd);
  groups_for(true, gr);
  bool rev = gr.size() < gd.size();
  auto& g = rev ? gr : gd;
  MapDev& M = C.maps[k];
  M.inv = invmask;
  M.rev = rev ? 1 : 0;
  M.ng = (int)g.size();
  M.goff = goff;
  for (size_t j = 0; j < g.size(); ++j) {
    C.gmask[goff + j] = g[j].second;
    C.gshl[goff + j] = g[j].first > 0 ? g[j].first : 0;
    C.gshr[goff + j] = g[j].first < 0 ? -g[j].first : 0;
  }
  goff += (int)g.size();
  // fermion re-ordering masks (mode j <-> bit N-1-j; mode order = site order)
  std::vector<int> newpos(N);
  for (int i = 0; i < N; ++i) newpos[perm[i]] = i;
  for (int b = 0; b < N; ++b) {
    int jb = N - 1 - b;
    uint64_t w = 0;
    for (int b2 = 0; b2 < N; ++b2) {
      int j2 = N - 1 - b2;
      if (j2 < jb && newpos[j2] > newpos[jb]) w |= 1ull << b2;
    }
    C.wmask[k * QSX_MAXN + b] = w;
  }
  return true;
}

int map_period(const int32_t* map, int N) {
  std::vector<int> perm(N), inv(N), sites(N), order(N);
  for (int i = 0; i < N; ++i) {
    int m = map[i];
    if (m < 0) { perm[i] = -(m + 1); inv[i] = 1; } else { perm[i] = m; inv[i] = 0; }
    sites[i] = i; order[i] = i;
  }
  int per = 0;
  while (per < 1 << 20) {
    std::vector<int> t = sites;
    for (int i = 0; i < N; ++i) if (inv[i]) t[i] = -(t[i] + 1);
    for (int i = 0; i < N; ++i) sites[i] = t[perm[i]];
    ++per;
    if (sites == order) break;
  }
  return per;
}

double2 exact_phase(int64_t num, int64_t den) {
  // exp(2 pi i num/den) with the quarter turns exact
  int64_t g = gcd64(num, den);
  if (g) { num /= g; den /= g; }
  num %= den; if (num < 0) num += den;
  if (num == 0) return make_double2(1.0, 0.0);
  if (2 * num == den) return make_double2(-1.0, 0.0);
  if (4 * num == den) return make_double2(0.0, 1.0);
  if (4 * num == 3 * den) return make_double2(0.0, -1.0);
  double a = 2.0 * M_PI * (double)num / (double)den;
  return make_double2(std::cos(a), std::sin(a));
}

void fill_binom(RankDev& R) {
  for (int n = 0; n <= 64; ++n)
    for (int k = 0; k <= 32; ++k) {
      if (k == 0) R.binom[n][k] = 1;
      else if (n == 0) R.binom[n][k] = 0;
      else {
        // C(n,k) = C(n-1,k-1) + C(n-1,k), saturate on overflow
        int64_t a = R.binom[n - 1][k - 1], b = (k <= 32) ? R.binom[n - 1][k] : 0;
        R.binom[n][k] = (a > INT64_MAX - b) ? INT64_MAX : a + b;
      }
    }
}

}  // namespace

extern "C" int qsx_core_create(int kind, int N, const int32_t* maps, const int32_t* pers,
                               const int32_t* qs, int n_maps, int double_occupancy,
                               qsx_core** out) {
  if (!out || N <= 0 || N > 64 || n_maps < 0 || n_maps > QSX_MAX_MAPS) {
    qsx_set_error("qsx_core_create: invalid N or number of maps (N <= 64, maps <= 32)");
    return QSX_ERR_ARG;
  }
  if (kind == KIND_SPINFUL && (N % 2)) {
    qsx_set_error("qsx_core_create: spinful cores take N = 2L bits");
    return QSX_ERR_ARG;
  }
  qsx_core* c = new qsx_core();
  std::memset(&c->h, 0, sizeof(CoreDev));
  std::memset(&c->rank_h, 0, sizeof(RankDev));
  c->h.N = N;
  c->h.kind = kind;
  c->h.nmaps = n_maps;
  c->h.L = (kind == KIND_SPINFUL) ? N / 2 : N;
  c->h.double_occ = double_occupancy ? 1 : 0;
  c->h.fermion = (kind != KIND_HCB) ? 1 : 0;
  int goff = 0;
  int64_t G = 1;
  for (int k = 0; k < n_maps; ++k) {
    if (!compile_map(maps + (size_t)k * N, N, c->h, k, goff)) {
      delete c;
      qsx_set_error("map must be a one-to-one site mapping.");
      return QSX_ERR_ARG;
    }
    int per = pers ? pers[k] : map_period(maps + (size_t)k * N, N);
    if (per != map_period(maps + (size_t)k * N, N)) {
      delete c;
      qsx_set_error("qsx_core_create: period does not match the map");
      return QSX_ERR_ARG;
    }
    c->h.maps[k].per = per;
    if (c->h.fermion && c->h.maps[k].inv) {
      delete c;
      qsx_set_error("fermion symmetry maps with population inversion are not supported");
      return QSX_ERR_UNSUPPORTED;
    }
    G *= per;
    if (G > (1 << 24)) {
      delete c;
      qsx_set_error("qsx_core_create: group order above 2^24");
      return QSX_ERR_UNSUPPORTED;
    }
  }
  c->h.G = (int)G;
  // characters chi(code) = exp(+2 pi i sum_k q_k l_k / per_k)
  c->phase_h.resize(G);
  int64_t lcm = 1;
  for (int k = 0; k < n_maps; ++k) lcm = lcm / gcd64(lcm, c->h.maps[k].per) * c->h.maps[k].per;
  for (int64_t code = 0; code < G; ++code) {
    int64_t rem = code, num = 0;
    for (int k = n_maps - 1; k >= 0; --k) {
      int per = c->h.maps[k].per;
      int64_t l = rem % per;
      rem /= per;
      int64_t q = qs ? qs[k] : 0;
      num += ((q % per + per) % per) * l * (lcm / per);
      num %= lcm;
    }
    c->phase_h[code] = exact_phase(num, lcm);
  }
  fill_binom(c->rank_h);
  c->rank_h.N = N;
  c->rank_h.L = c->h.L;
  hipError_t e = hipMalloc(&c->d, sizeof(CoreDev));
  if (e == hipSuccess) e = hipMemcpy(c->d, &c->h, sizeof(CoreDev), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->phase, sizeof(double2) * G);
  if (e == hipSuccess) e = hipMemcpy(c->phase, c->phase_h.data(), sizeof(double2) * G, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->rank_d, sizeof(RankDev));
  if (e == hipSuccess) e = hipMemcpy(c->rank_d, &c->rank_h, sizeof(RankDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    qsx_set_error(std::string("qsx_core_create: ") + hipGetErrorString(e));
    if (c->d) (void)hipFree(c->d);
    if (c->phase) (void)hipFree(c->phase);
    if (c->rank_d) (void)hipFree(c->rank_d);
    delete c;
    return QSX_ERR_HIP;
  }
  *out = c;
  return QSX_OK;
}

extern "C" void qsx_core_destroy(qsx_core* c) {
  if (!c) return;
  if (c->d) (void)hipFree(c->d);
  if (c->phase) (void)hipFree(c->phase);
  if (c->rank_d) (void)hipFree(c->rank_d);
  delete c;
}

extern "C" int qsx_core_group_order(const qsx_core* c) { return c ? c->h.G : -1; }

// Debug/introspection: compiled groups per map (for host tests).
extern "C" int qsx_core_map_info(const qsx_core* c, int k, int* rev, int* ngroups, int* per) {
  if (!c || k < 0 || k >= c->h.nmaps) return QSX_ERR_ARG;
  *rev = c->h.maps[k].rev;
  *ngroups = c->h.maps[k].ng;
  *per = c->h.maps[k].per;
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// device: candidate enumeration
// ---------------------------------------------------------------------------
struct EnumDev {
  int32_t mode;      // 0: single hcb sector (Np), 1: 2^N with popcount filter mask,
                     // 2: spinful single (Nup,Ndn), 3: 4^L with (Nup,Ndn) pair filter
  int32_t Np, Nup, Ndn;
  int64_t total;
  uint64_t popmask;  // mode 1: bit k set if popcount k allowed (all set = no pcon)
  uint64_t pairmask[65];  // mode 3: pairmask[nu] bit nd set if (nu, nd) allowed
};

template <typename I>
__device__ __forceinline__ I unrank_comb(const RankDev* __restrict__ R, int64_t r, int k, int N) {
  I s = 0;
  for (int j = k; j >= 1; --j) {
    int p = j - 1;
    while (p + 1 < N && R->binom[p + 1][j] <= r) ++p;
    s |= ((I)1) << p;
    r -= R->binom[p][j];
  }
  return s;
}

template <typename I>
__device__ __forceinline__ I gosper_next(I s) {
  I c = s & (~s + 1);
  I r = s + c;
  int tz = (sizeof(I) == 8) ? __builtin_ctzll((uint64_t)c) : __builtin_ctz((uint32_t)c);
  return (((r ^ s) >> 2) >> tz) | r;
}

// Iterate candidates [start, start+cnt) in ascending order calling f(s).
template <typename I, typename F>
__device__ __forceinline__ void for_candidates(const EnumDev* __restrict__ E, const RankDev* __restrict__ R,
                                               int N, int L, int double_occ, int64_t start, int64_t cnt, F&& f) {
  if (cnt <= 0) return;
  if (E->mode == 0) {
    if (E->Np == 0) { f((I)0); return; }
    I s = unrank_comb<I>(R, start, E->Np, N);
    for (int64_t i = 0; i < cnt; ++i) {
      f(s);
      if (i + 1 < cnt) s = gosper_next<I>(s);
    }
  } else if (E->mode == 1) {
    for (int64_t i = 0; i < cnt; ++i) {
      I s = (I)(start + i);
      if ((E->popmask >> popc<I>(s)) & 1ull) {
        if (!double_occ) {
          const I mlo = (((I)1) << L) - 1;
          if ((s >> L) & s & mlo) continue;
        }
        f(s);
      }
    }
  } else if (E->mode == 2) {
    const int64_t cdn = R->binom[L][E->Ndn];
    int64_t iu = start / cdn, id = start % cdn;
    I up = (E->Nup == 0) ? (I)0 : unrank_comb<I>(R, iu, E->Nup, L);
    I dn = (E->Ndn == 0) ? (I)0 : unrank_comb<I>(R, id, E->Ndn, L);
    for (int64_t i = 0; i < cnt; ++i) {
      if (double_occ || !(up & dn)) f((I)((up << L) | dn));
      if (i + 1 < cnt) {
        ++id;
        if (id == cdn) {
          id = 0;
          ++iu;
          up = (E->Nup == 0) ? (I)0 : gosper_next<I>(up);
          dn = (E->Ndn == 0) ? (I)0 : unrank_comb<I>(R, 0, E->Ndn, L);
        } else {
          dn = gosper_next<I>(dn);
        }
      }
    }
  } else {
    const I mlo = (((I)1) << L) - 1;
    for (int64_t i = 0; i < cnt; ++i) {
      I s = (I)(start + i);
      I up = s >> L, dn = s & mlo;
      if (!((E->pairmask[popc<I>(up)] >> popc<I>(dn)) & 1ull)) continue;
      if (!double_occ && (up & dn)) continue;
      f(s);
    }
  }
}

// s is kept when it is the orbit maximum and its character sum over the
// stabiliser is positive; returns n (0 if rejected).  Early exit on the first
// image above s.
template <typename I, int M, bool FERM>
__device__ __forceinline__ int check_state(const CoreDev* __restrict__ C, const double2* __restrict__ ph, I s) {
  double nacc = 0.0;
  auto vis = [&](I t, int code, int sign) -> bool {
    if (t > s) return false;
    if (t == s) nacc += sign * ph[code].x;
    return true;
  };
  if (!visit_orbit<I, M, FERM>(C, s, vis)) return 0;
  int n = (int)rint(nacc);
  return n > 0 ? n : 0;
}

template <typename I, int M, bool FERM>
__global__ void __launch_bounds__(256) k_count_reps(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                                    const EnumDev* __restrict__ E, const RankDev* __restrict__ R,
                                                    int64_t chunk, int64_t nchunks, int64_t* __restrict__ counts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nchunks) return;
  const int64_t start = t * chunk;
  const int64_t cnt = min(chunk, E->total - start);
  int64_t k = 0;
  for_candidates<I>(E, R, C->N, C->L, C->double_occ, start, cnt, [&](I s) {
    if (check_state<I, M, FERM>(C, ph, s) > 0) ++k;
  });
  counts[t] = k;
}

template <typename I, int M, bool FERM>
__global__ void __launch_bounds__(256) k_write_reps(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                                    const EnumDev* __restrict__ E, const RankDev* __restrict__ R,
                                                    int64_t chunk, int64_t nchunks, const int64_t* __restrict__ offs,
                                                    int64_t Ns, I* __restrict__ basis, void* __restrict__ nout,
                                                    int n_bytes) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nchunks) return;
  const int64_t start = t * chunk;
  const int64_t cnt = min(chunk, E->total - start);
  int64_t pos = offs[t];
  for_candidates<I>(E, R, C->N, C->L, C->double_occ, start, cnt, [&](I s) {
    int n = check_state<I, M, FERM>(C, ph, s);
    if (n > 0) {
      const int64_t o = Ns - 1 - pos;  // ascending enumeration -> descending storage
      basis[o] = s;
      if (n_bytes == 1) ((uint8_t*)nout)[o] = (uint8_t)n;
      else if (n_bytes == 2) ((uint16_t*)nout)[o] = (uint16_t)n;
      else ((uint32_t*)nout)[o] = (uint32_t)n;
      ++pos;
    }
  });
}

#define QSX_M_SWITCH(nm, BODY)          \
  switch (nm) {                         \
    case 0: { constexpr int MM = 0; BODY; } break; \
    case 1: { constexpr int MM = 1; BODY; } break; \
    case 2: { constexpr int MM = 2; BODY; } break; \
    case 3: { constexpr int MM = 3; BODY; } break; \
    case 4: { constexpr int MM = 4; BODY; } break; \
    case 5: { constexpr int MM = 5; BODY; } break; \
    default: { constexpr int MM = -1; BODY; } break; \
  }

template <typename I, int M, bool FERM>
static int make_basis_impl(qsx_core* c, const EnumDev& Eh, void* basis_out, void* n_out, int n_bytes,
                           int64_t capacity, int64_t* Ns_out, hipStream_t st) {
  const int64_t chunk = 64;
  const int64_t nchunks = (Eh.total + chunk - 1) / chunk;
  EnumDev* Ed = nullptr;
  int64_t* counts = nullptr;
  int64_t* offs = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = QSX_OK;
  int64_t last_cnt = 0, last_off = 0, Ns = 0;
  const int blk = 256;
  if (nchunks == 0) { *Ns_out = 0; return QSX_OK; }
  QSX_CHECK_HIP(hipMalloc(&Ed, sizeof(EnumDev)));
  QSX_CHECK_HIP(hipMemcpyAsync(Ed, &Eh, sizeof(EnumDev), hipMemcpyHostToDevice, st));
  QSX_CHECK_HIP(hipMalloc(&counts, sizeof(int64_t) * nchunks));
  QSX_CHECK_HIP(hipMalloc(&offs, sizeof(int64_t) * nchunks));
  hipLaunchKernelGGL((k_count_reps<I, M, FERM>), dim3(qsx_grid(nchunks, blk)), dim3(blk), 0, st, c->d, c->phase,
                     Ed, c->rank_d, chunk, nchunks, counts);
  QSX_LAUNCH_CHECK();
  QSX_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, offs, (int)nchunks, st));
  QSX_CHECK_HIP(hipMalloc(&tmp, tmp_bytes + 16));
  QSX_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, offs, (int)nchunks, st));
  QSX_CHECK_HIP(hipMemcpyAsync(&last_cnt, counts + nchunks - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipMemcpyAsync(&last_off, offs + nchunks - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  QSX_CHECK_HIP(hipStreamSynchronize(st));
  Ns = last_cnt + last_off;
  *Ns_out = Ns;
  if (basis_out && Ns > capacity) {
    rc = QSX_ERR_OVERFLOW;
  } else if (basis_out && Ns > 0) {
    hipLaunchKernelGGL((k_write_reps<I, M, FERM>), dim3(qsx_grid(nchunks, blk)), dim3(blk), 0, st, c->d,
                       c->phase, Ed, c->rank_d, chunk, nchunks, offs, Ns, (I*)basis_out, n_out, n_bytes);
    QSX_LAUNCH_CHECK();
    QSX_CHECK_HIP(hipStreamSynchronize(st));
  }
  (void)hipFree(tmp);
  (void)hipFree(counts);
  (void)hipFree(offs);
  (void)hipFree(Ed);
  return rc;
}

static int64_t binom_host(int n, int k) {
  if (k < 0 || k > n) return 0;
  long double r = 1;
  int64_t v = 1;
  for (int i = 1; i <= k; ++i) { v = v * (n - k + i) / i; (void)r; }
  return v;
}

// Np: hcb -> list of particle numbers (n_Np entries, n_Np = 0 for no
// conservation); spinful -> n_Np pairs (Nup, Ndown) flattened.
// basis_out == NULL: count only.  *Ns_out = Ns; QSX_ERR_OVERFLOW when
// capacity < Ns (then nothing is written and *Ns_out is the size needed).
extern "C" int qsx_make_basis(qsx_core* c, const int32_t* Np, int n_Np, void* basis_out, void* n_out, int n_bytes,
                              int64_t capacity, int64_t* Ns_out, void* stream) {
  if (!c || !Ns_out) { qsx_set_error("qsx_make_basis: null core or Ns_out"); return QSX_ERR_ARG; }
  *Ns_out = -1;
  hipStream_t st = (hipStream_t)stream;
  const int N = c->h.N, L = c->h.L;
  if (n_bytes != 1 && n_bytes != 2 && n_bytes != 4) { qsx_set_error("n_bytes must be 1, 2 or 4"); return QSX_ERR_ARG; }
  EnumDev E;
  std::memset(&E, 0, sizeof(E));
  RankDev& R = c->rank_h;
  c->rank_ok = 0;
  c->sector_ok = 0;
  if (c->h.kind != KIND_SPINFUL) {
    if (n_Np == 1 && Np[0] >= 0 && Np[0] <= N && Np[0] <= 32) {
      E.mode = 0;
      E.Np = Np[0];
      E.total = binom_host(N, Np[0]);
      R.mode = 0; R.Np = Np[0]; R.Ns = E.total;
      c->sector_ok = 1; c->sector_total = E.total;
      if (c->h.nmaps == 0) c->rank_ok = 1;
    } else {
      if (N > 40) { qsx_set_error("qsx_make_basis: several or no particle sectors need N <= 40"); return QSX_ERR_UNSUPPORTED; }
      E.mode = 1;
      E.total = (int64_t)1 << N;
      if (n_Np == 0) E.popmask = ~0ull;
      for (int i = 0; i < n_Np; ++i) {
        if (Np[i] < 0 || Np[i] > N) { qsx_set_error("particle number Nb must satisfy: 0 <= Nb <= N"); return QSX_ERR_ARG; }
        E.popmask |= 1ull << Np[i];
      }
      if (n_Np == 0 && N <= 32) {
        R.mode = 2; R.Ns = E.total;
        c->sector_ok = 1; c->sector_total = E.total;
        if (c->h.nmaps == 0) c->rank_ok = 1;
      }
    }
  } else {
    if (n_Np == 1) {
      E.mode = 2;
      E.Nup = Np[0]; E.Ndn = Np[1];
      if (E.Nup < 0 || E.Nup > L || E.Ndn < 0 || E.Ndn > L) { qsx_set_error("particle numbers in Nf must satisfy: 0 <= n <= N"); return QSX_ERR_ARG; }
      E.total = binom_host(L, E.Nup) * binom_host(L, E.Ndn);
      R.mode = 1; R.Nup = E.Nup; R.Ndn = E.Ndn; R.Ns = E.total; R.cdn = binom_host(L, E.Ndn);
      c->sector_ok = 1; c->sector_total = E.total;
      if (c->h.nmaps == 0 && c->h.double_occ) c->rank_ok = 1;
    } else {
      if (N > 40) { qsx_set_error("qsx_make_basis: several or no particle sectors need 2L <= 40"); return QSX_ERR_UNSUPPORTED; }
      E.mode = 3;
      E.total = (int64_t)1 << N;
      for (int i = 0; i <= 64; ++i) E.pairmask[i] = (n_Np == 0) ? ~0ull : 0ull;
      for (int i = 0; i < n_Np; ++i) {
        int nu = Np[2 * i], nd = Np[2 * i + 1];
        if (nu < 0 || nu > L || nd < 0 || nd > L) { qsx_set_error("particle numbers in Nf must satisfy: 0 <= n <= N"); return QSX_ERR_ARG; }
        E.pairmask[nu] |= 1ull << nd;
      }
    }
  }
  QSX_CHECK_HIP(hipMemcpy(c->rank_d, &R, sizeof(RankDev), hipMemcpyHostToDevice));
  int64_t Ns = 0;
  int rc = QSX_OK;
  const bool ferm = c->h.fermion != 0;
  if (N <= 32) {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, rc = (make_basis_impl<uint32_t, MM, true>(c, E, basis_out, n_out, n_bytes, capacity, &Ns, st))); }
    else { QSX_M_SWITCH(c->h.nmaps, rc = (make_basis_impl<uint32_t, MM, false>(c, E, basis_out, n_out, n_bytes, capacity, &Ns, st))); }
  } else {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, rc = (make_basis_impl<uint64_t, MM, true>(c, E, basis_out, n_out, n_bytes, capacity, &Ns, st))); }
    else { QSX_M_SWITCH(c->h.nmaps, rc = (make_basis_impl<uint64_t, MM, false>(c, E, basis_out, n_out, n_bytes, capacity, &Ns, st))); }
  }
  if (rc == QSX_OK || rc == QSX_ERR_OVERFLOW) *Ns_out = Ns;
  if (rc == QSX_ERR_OVERFLOW) qsx_set_error("qsx_make_basis: capacity below the number of states");
  return rc;
}

// ---------------------------------------------------------------------------
// prefix blocks: begin[p], end[p] bound the states whose top N_p bits equal p
// ---------------------------------------------------------------------------
template <typename I>
__device__ __forceinline__ int64_t count_ge(const I* __restrict__ b, int64_t Ns, uint64_t key, bool overflow) {
  if (overflow) return 0;
  int64_t lo = 0, hi = Ns;  // b descending: first index with b[i] < key
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((uint64_t)b[mid] >= key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <typename I>
__global__ void k_blocks(const I* __restrict__ b, int64_t Ns, int N, int Np, int64_t* __restrict__ begin,
                         int64_t* __restrict__ end) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = (int64_t)1 << Np;
  if (p >= P) return;
  const int sh = N - Np;
  const uint64_t k0 = (uint64_t)p << sh;
  const bool ovf = (p + 1 == P) && (sh + Np >= 64);
  const uint64_t k1 = ovf ? 0 : ((uint64_t)(p + 1) << sh);
  end[p] = count_ge<I>(b, Ns, k0, false);
  begin[p] = count_ge<I>(b, Ns, k1, ovf);
}

extern "C" int qsx_make_basis_blocks(qsx_core* c, int N_p, const void* basis, int64_t Ns, int64_t* begin,
                                     int64_t* end, void* stream) {
  if (!c || N_p < 0 || N_p > c->h.N || N_p > 40) { qsx_set_error("qsx_make_basis_blocks: bad N_p"); return QSX_ERR_ARG; }
  if (N_p == 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t P = (int64_t)1 << N_p;
  if (c->h.N <= 32)
    hipLaunchKernelGGL(k_blocks<uint32_t>, dim3(qsx_grid(P, 256)), dim3(256), 0, st, (const uint32_t*)basis, Ns,
                       c->h.N, N_p, begin, end);
  else
    hipLaunchKernelGGL(k_blocks<uint64_t>, dim3(qsx_grid(P, 256)), dim3(256), 0, st, (const uint64_t*)basis, Ns,
                       c->h.N, N_p, begin, end);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// representative / normalization of arbitrary states
// ---------------------------------------------------------------------------
template <typename I, int M, bool FERM>
__global__ void k_representative(const CoreDev* __restrict__ C, const I* __restrict__ states, I* __restrict__ out,
                                 int32_t* __restrict__ g_out, int8_t* __restrict__ sign_out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  OrbitRes R;
  const I r = orbit_max<I, M, FERM>(C, states[i], R);
  out[i] = r;
  if (g_out) {
    int code = R.code;
    for (int k = C->nmaps - 1; k >= 0; --k) {
      const int per = C->maps[k].per;
      g_out[i * C->nmaps + k] = code % per;
      code /= per;
    }
  }
  if (sign_out) sign_out[i] = (int8_t)R.sign;
}

template <typename I, int M, bool FERM>
__global__ void k_normalization(const CoreDev* __restrict__ C, const double2* __restrict__ ph,
                                const I* __restrict__ states, void* __restrict__ out, int out_bytes, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const I s = states[i];
  double nacc = 0.0;
  auto vis = [&](I t, int code, int sign) -> bool {
    if (t == s) nacc += sign * ph[code].x;
    return true;
  };
  visit_orbit<I, M, FERM>(C, s, vis);
  int64_t v = (int64_t)rint(nacc);
  if (v < 0) v = 0;
  if (out_bytes == 1) ((uint8_t*)out)[i] = (uint8_t)v;
  else if (out_bytes == 2) ((uint16_t*)out)[i] = (uint16_t)v;
  else if (out_bytes == 4) ((uint32_t*)out)[i] = (uint32_t)v;
  else ((uint64_t*)out)[i] = (uint64_t)v;
}

extern "C" int qsx_representative(qsx_core* c, const void* states, void* out, int32_t* g_out, int8_t* sign_out,
                                  int64_t n, void* stream) {
  if (!c) { qsx_set_error("qsx_representative: null core"); return QSX_ERR_ARG; }
  if (n <= 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool ferm = c->h.fermion != 0;
  const int blk = 256;
  if (c->h.N <= 32) {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_representative<uint32_t, MM, true>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, (const uint32_t*)states, (uint32_t*)out, g_out, sign_out, n)); }
    else { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_representative<uint32_t, MM, false>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, (const uint32_t*)states, (uint32_t*)out, g_out, sign_out, n)); }
  } else {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_representative<uint64_t, MM, true>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, (const uint64_t*)states, (uint64_t*)out, g_out, sign_out, n)); }
    else { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_representative<uint64_t, MM, false>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, (const uint64_t*)states, (uint64_t*)out, g_out, sign_out, n)); }
  }
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_normalization(qsx_core* c, const void* states, void* out, int out_bytes, int64_t n,
                                 void* stream) {
  if (!c || (out_bytes != 1 && out_bytes != 2 && out_bytes != 4 && out_bytes != 8)) {
    qsx_set_error("qsx_normalization: bad arguments");
    return QSX_ERR_ARG;
  }
  if (n <= 0) return QSX_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool ferm = c->h.fermion != 0;
  const int blk = 256;
  if (c->h.N <= 32) {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_normalization<uint32_t, MM, true>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, c->phase, (const uint32_t*)states, out, out_bytes, n)); }
    else { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_normalization<uint32_t, MM, false>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, c->phase, (const uint32_t*)states, out, out_bytes, n)); }
  } else {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_normalization<uint64_t, MM, true>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, c->phase, (const uint64_t*)states, out, out_bytes, n)); }
    else { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_normalization<uint64_t, MM, false>), dim3(qsx_grid(n, blk)), dim3(blk), 0, st, c->d, c->phase, (const uint64_t*)states, out, out_bytes, n)); }
  }
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// ---------------------------------------------------------------------------
// sector lookup table: for every state of the enumerated particle sector (in
// rank order) the basis index of its representative and the code of the
// group element reaching it; absent states hold all ones.
// ---------------------------------------------------------------------------
template <typename I>
__device__ __forceinline__ I sector_state(const RankDev* __restrict__ R, int64_t rk) {
  if (R->mode == 0) return R->Np == 0 ? (I)0 : unrank_comb<I>(R, rk, R->Np, R->N);
  if (R->mode == 1) {
    const int64_t iu = rk / R->cdn, id = rk % R->cdn;
    const I up = R->Nup == 0 ? (I)0 : unrank_comb<I>(R, iu, R->Nup, R->L);
    const I dn = R->Ndn == 0 ? (I)0 : unrank_comb<I>(R, id, R->Ndn, R->L);
    return (up << R->L) | dn;
  }
  return (I)rk;
}

template <typename I>
__device__ __forceinline__ I sector_next(const RankDev* __restrict__ R, I s, int64_t rk_next) {
  if (R->mode == 0) return R->Np == 0 ? s : gosper_next<I>(s);
  if (R->mode == 1) {
    if (rk_next % R->cdn == 0) return sector_state<I>(R, rk_next);
    const I mlo = (((I)1) << R->L) - 1;
    return (s & ~mlo) | gosper_next<I>(s & mlo);
  }
  return s + 1;
}

// Coefficient classes of the table: the distinct normalisations n > 0 of
// the basis (sorted), at most QSX_TABLE_NCLS of them.
constexpr int QSX_TABLE_NCLS = 64;

__device__ __forceinline__ uint32_t table_n(const void* p, int nb, int64_t i) {
  return nb == 1 ? ((const uint8_t*)p)[i] : nb == 2 ? ((const uint16_t*)p)[i] : ((const uint32_t*)p)[i];
}

// ncls > 0: the code field holds chi_cls[code] + nchi * class(n_c), an id
// into the coefficient table chi * sqrt(n_c) whose rows are the distinct
// characters (chi_cls maps each group element to its character's row);
// ncls = 0: the group element code.
template <typename I, int M, bool FERM>
__global__ void __launch_bounds__(256) k_table(const CoreDev* __restrict__ C, const RankDev* __restrict__ R,
                                               const I* __restrict__ basis, int64_t Ns, const int64_t* __restrict__ begin,
                                               const int64_t* __restrict__ end, int N_p, int64_t total, int64_t chunk,
                                               void* __restrict__ table, int u64, int idx_bits, const void* nrm,
                                               int n_bytes, const uint32_t* __restrict__ nvals, int ncls,
                                               const int32_t* __restrict__ chi_cls, int nchi) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t * chunk;
  if (a >= total) return;
  const int64_t b = min(total, a + chunk);
  I s = sector_state<I>(R, a);
  for (int64_t rk = a; rk < b; ++rk) {
    OrbitRes O;
    const I rep = orbit_max<I, M, FERM>(C, s, O);
    int64_t lo = 0, hi = Ns;
    if (N_p > 0) {
      const int64_t p = (int64_t)((uint64_t)rep >> (C->N - N_p));
      lo = begin[p];
      hi = end[p];
    }
    const int64_t idx = bsearch_desc<I>(basis, lo, hi, rep);
    uint32_t code = (uint32_t)O.code;
    if (ncls > 0 && idx >= 0) {
      const uint32_t n = table_n(nrm, n_bytes, idx);
      int k = 0;
      while (k + 1 < ncls && nvals[k] != n) ++k;
      code = (uint32_t)chi_cls[O.code] + (uint32_t)nchi * (uint32_t)k;
    }
    if (u64) {
      uint64_t e = ~0ull;
      if (idx >= 0) e = (uint64_t)idx | ((uint64_t)code << 40) | (O.sign < 0 ? (1ull << 63) : 0ull);
      ((uint64_t*)table)[rk] = e;
    } else {
      uint32_t e = 0xFFFFFFFFu;
      if (idx >= 0) e = (uint32_t)idx | (code << idx_bits);
      ((uint32_t*)table)[rk] = e;
    }
    if (rk + 1 < b) s = sector_next<I>(R, s, rk + 1);
  }
}

// Distinct normalisations n > 0 of the basis (sorted); false when there are
// more than QSX_TABLE_NCLS.
static int table_classes(const void* nrm, int n_bytes, int64_t Ns, std::vector<uint32_t>& vals, bool& fits) {
  fits = false;
  vals.clear();
  if (!nrm || Ns <= 0) return QSX_OK;
  std::vector<unsigned char> h((size_t)Ns * n_bytes);
  QSX_CHECK_HIP(hipMemcpy(h.data(), nrm, h.size(), hipMemcpyDeviceToHost));
  std::vector<uint32_t> v;
  for (int64_t i = 0; i < Ns; ++i) {
    uint32_t n = n_bytes == 1 ? h[i] : n_bytes == 2 ? ((const uint16_t*)h.data())[i] : ((const uint32_t*)h.data())[i];
    if (n == 0) continue;
    if (std::find(v.begin(), v.end(), n) == v.end()) {
      v.push_back(n);
      if ((int)v.size() > QSX_TABLE_NCLS) return QSX_OK;
    }
  }
  std::sort(v.begin(), v.end());
  vals = v;
  fits = true;
  return QSX_OK;
}

// Distinct characters chi(g) of the group (bitwise distinct, in order of
// first appearance) and the row of each group element among them.  Elements
// with equal characters give equal matrix elements, so the coefficient table
// needs one row per distinct character, not per element: for k = 0 with
// +1 parity and inversion eigenvalues that is a single row.
static void chi_classes(const qsx_core* c, std::vector<int32_t>& cls, std::vector<double2>& vals) {
  cls.assign(c->h.G, 0);
  vals.clear();
  for (int g = 0; g < c->h.G; ++g) {
    const double2 p = c->phase_h[g];
    int k = 0;
    while (k < (int)vals.size() && std::memcmp(&vals[k], &p, sizeof(double2)) != 0) ++k;
    if (k == (int)vals.size()) vals.push_back(p);
    cls[g] = k;
  }
}

// Size and layout of the lookup table for the sector of the last make_basis:
// entries, bytes per entry (4 or 8), index bits of the 4-byte layout, and the
// length of the coefficient table (0: entries hold the group element; the
// kernels then read n_c of the target).  nrm / n_bytes: the basis n array.
// Returns QSX_ERR_UNSUPPORTED when the basis was not built from one sector.
extern "C" int qsx_table_info(const qsx_core* c, int64_t Ns, const void* nrm, int n_bytes, int64_t* entries,
                              int* entry_bytes, int* idx_bits, int64_t* lut_len) {
  if (!c || !c->sector_ok || c->h.nmaps == 0) {
    qsx_set_error("lookup table needs a symmetric basis made from one particle sector");
    return QSX_ERR_UNSUPPORTED;
  }
  if (!entries || !entry_bytes || !idx_bits || !lut_len || (nrm && n_bytes != 1 && n_bytes != 2 && n_bytes != 4)) {
    qsx_set_error("qsx_table_info: bad arguments");
    return QSX_ERR_ARG;
  }
  int ib = 1;
  while (ib < 40 && ((int64_t)1 << ib) <= Ns) ++ib;
  int cb = 0;
  while ((1 << cb) < c->h.G) ++cb;
  std::vector<uint32_t> vals;
  bool fits = false;
  const int rc = table_classes(nrm, n_bytes, Ns, vals, fits);
  if (rc) return rc;
  std::vector<int32_t> gcls;
  std::vector<double2> chis;
  chi_classes(c, gcls, chis);
  const int64_t nl = fits ? (int64_t)chis.size() * (int64_t)vals.size() : 0;
  int lb = 0;
  while (nl > 0 && ((int64_t)1 << lb) < nl) ++lb;
  *entries = c->sector_total;
  *idx_bits = ib;
  if (!c->h.fermion && nl > 0 && ib + lb <= 31) {
    *entry_bytes = 4;
    *lut_len = nl;
  } else if (!c->h.fermion && ib + cb <= 31) {
    *entry_bytes = 4;
    *lut_len = 0;
  } else {
    *entry_bytes = 8;
    *lut_len = (nl > 0 && nl <= (1 << 23)) ? nl : 0;
  }
  if (*entry_bytes == 8 && (Ns >= (1ll << 40) || c->h.G > (1 << 23))) {
    qsx_set_error("lookup table: basis or group too large");
    return QSX_ERR_UNSUPPORTED;
  }
  return QSX_OK;
}

// Fills table (entries as described by qsx_table_info) and, when lut_len > 0,
// the coefficient table lut[r + nchi * class] = chi_r * sqrt(n_class) over
// the distinct characters chi_r (lut_len double2 of device memory).
// Synchronous.
extern "C" int qsx_table_build(qsx_core* c, const void* basis, int64_t Ns, const int64_t* begin, const int64_t* end,
                               int N_p, const void* nrm, int n_bytes, void* table, int entry_bytes, int idx_bits,
                               void* lut, int64_t lut_len, void* stream) {
  if (!c || !c->sector_ok || (entry_bytes != 4 && entry_bytes != 8) || (lut_len > 0 && (!lut || !nrm))) {
    qsx_set_error("qsx_table_build: bad arguments");
    return QSX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  std::vector<uint32_t> vals;
  uint32_t* nv_d = nullptr;
  int32_t* cls_d = nullptr;
  int ncls = 0, nchi = 0;
  if (lut_len > 0) {
    bool fits = false;
    const int rc = table_classes(nrm, n_bytes, Ns, vals, fits);
    if (rc) return rc;
    std::vector<int32_t> gcls;
    std::vector<double2> chis;
    chi_classes(c, gcls, chis);
    nchi = (int)chis.size();
    if (!fits || (int64_t)nchi * (int64_t)vals.size() != lut_len) {
      qsx_set_error("qsx_table_build: coefficient table length does not match the basis");
      return QSX_ERR_ARG;
    }
    ncls = (int)vals.size();
    std::vector<double2> lh((size_t)lut_len);
    for (int k = 0; k < ncls; ++k) {
      const double r = std::sqrt((double)vals[k]);
      for (int j = 0; j < nchi; ++j)
        lh[(size_t)j + (size_t)nchi * k] = make_double2(r * chis[j].x, r * chis[j].y);
    }
    QSX_CHECK_HIP(hipMemcpy(lut, lh.data(), sizeof(double2) * lut_len, hipMemcpyHostToDevice));
    QSX_CHECK_HIP(hipMalloc(&nv_d, sizeof(uint32_t) * ncls));
    QSX_CHECK_HIP(hipMemcpy(nv_d, vals.data(), sizeof(uint32_t) * ncls, hipMemcpyHostToDevice));
    const hipError_t ec = hipMalloc(&cls_d, sizeof(int32_t) * gcls.size());
    if (ec != hipSuccess) {
      (void)hipFree(nv_d);
      qsx_set_error(std::string("hipMalloc: ") + hipGetErrorString(ec));
      return QSX_ERR_HIP;
    }
    QSX_CHECK_HIP(hipMemcpy(cls_d, gcls.data(), sizeof(int32_t) * gcls.size(), hipMemcpyHostToDevice));
  }
  const int64_t total = c->sector_total;
  const int64_t chunk = 16;
  const int64_t nth = (total + chunk - 1) / chunk;
  const dim3 grid((unsigned)qsx_grid(nth, 256)), blk(256);
  const int u64 = entry_bytes == 8;
  const bool ferm = c->h.fermion != 0;
  if (c->h.N <= 32) {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_table<uint32_t, MM, true>), grid, blk, 0, st, c->d, c->rank_d, (const uint32_t*)basis, Ns, begin, end, N_p, total, chunk, table, u64, idx_bits, nrm, n_bytes, nv_d, ncls, cls_d, nchi)); }
    else { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_table<uint32_t, MM, false>), grid, blk, 0, st, c->d, c->rank_d, (const uint32_t*)basis, Ns, begin, end, N_p, total, chunk, table, u64, idx_bits, nrm, n_bytes, nv_d, ncls, cls_d, nchi)); }
  } else {
    if (ferm) { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_table<uint64_t, MM, true>), grid, blk, 0, st, c->d, c->rank_d, (const uint64_t*)basis, Ns, begin, end, N_p, total, chunk, table, u64, idx_bits, nrm, n_bytes, nv_d, ncls, cls_d, nchi)); }
    else { QSX_M_SWITCH(c->h.nmaps, hipLaunchKernelGGL((k_table<uint64_t, MM, false>), grid, blk, 0, st, c->d, c->rank_d, (const uint64_t*)basis, Ns, begin, end, N_p, total, chunk, table, u64, idx_bits, nrm, n_bytes, nv_d, ncls, cls_d, nchi)); }
  }
  const hipError_t le = hipGetLastError();
  if (nv_d) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(nv_d);
    (void)hipFree(cls_d);
  }
  if (le != hipSuccess) {
    qsx_set_error(std::string("kernel launch: ") + hipGetErrorString(le));
    return QSX_ERR_HIP;
  }
  return QSX_OK;
}
