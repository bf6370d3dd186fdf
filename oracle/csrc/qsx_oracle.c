/* C/OpenMP restatement of the reference's symmetry-reduced basis and H.v
 * loops -- TEST INFRASTRUCTURE ONLY (the oracle).  Used by tests/ as an
 * independent checker at sizes numpy cannot reach, and by the bench's
 * cpu_baseline leg as the timed CPU path.  Never linked by the product.
 *
 * Independent of the device code: maps are applied site by site from the
 * integer map arrays (pull action, site j <-> bit N-1-j,
 * sphinx/source/tutorials/user_basis.rst:56-79), the group is enumerated by
 * nested powers of the maps, representatives are orbit maxima
 * (base_general.py:817-899 semantics), n = character sum over the stabiliser,
 * and the reduced element is J m chi(g) sqrt(n_r'/n_r) (hcb spin-1/2,
 * base_general.py:354-410 computed per term, as the reference's one pass per
 * term; a fused all-terms variant is timed as well).
 * Pins: tests/test_oracle_pins.py (test_c_oracle_matches_numpy_oracle) checks
 * this file against the numpy oracle
 * (brute force, projector identity) on small chains.
 */
#include <math.h>
#include <stdint.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QO_MAXM 16
#define QO_MAXN 64
#define QO_MAXOPS 8

typedef struct {
  int N, nmaps;
  int perm[QO_MAXM][QO_MAXN];
  int inv[QO_MAXM][QO_MAXN];
  int per[QO_MAXM];
  int q[QO_MAXM];
  int G;
  double *chr, *chi;  /* character per element code (map 0 most significant) */
} qo_group;

static uint64_t qo_apply(const qo_group* g, int k, uint64_t s) {
  uint64_t t = 0;
  const int N = g->N;
  for (int i = 0; i < N; ++i) {
    uint64_t b = (s >> (N - 1 - g->perm[k][i])) & 1ull;
    b ^= (uint64_t)g->inv[k][i];
    t |= b << (N - 1 - i);
  }
  return t;
}

qo_group* qo_group_create(int N, int nmaps, const int32_t* maps, const int32_t* pers, const int32_t* qs) {
  if (nmaps > QO_MAXM || N > QO_MAXN) return NULL;
  qo_group* g = (qo_group*)calloc(1, sizeof(qo_group));
  g->N = N;
  g->nmaps = nmaps;
  g->G = 1;
  for (int k = 0; k < nmaps; ++k) {
    for (int i = 0; i < N; ++i) {
      int m = maps[k * N + i];
      g->inv[k][i] = m < 0;
      g->perm[k][i] = m < 0 ? -(m + 1) : m;
    }
    g->per[k] = pers[k];
    g->q[k] = qs[k];
    g->G *= pers[k];
  }
  g->chr = (double*)malloc(sizeof(double) * g->G);
  g->chi = (double*)malloc(sizeof(double) * g->G);
  for (int c = 0; c < g->G; ++c) {
    int rem = c;
    double ph = 0.0;
    for (int k = nmaps - 1; k >= 0; --k) {
      int l = rem % g->per[k];
      rem /= g->per[k];
      ph += (double)g->q[k] * l / g->per[k];
    }
    ph -= floor(ph);
    double re = cos(2 * M_PI * ph), im = sin(2 * M_PI * ph);
    if (fabs(ph) < 1e-15) { re = 1; im = 0; }
    else if (fabs(ph - 0.5) < 1e-15) { re = -1; im = 0; }
    else if (fabs(ph - 0.25) < 1e-15) { re = 0; im = 1; }
    else if (fabs(ph - 0.75) < 1e-15) { re = 0; im = -1; }
    g->chr[c] = re;
    g->chi[c] = im;
  }
  return g;
}

void qo_group_destroy(qo_group* g) {
  if (!g) return;
  free(g->chr);
  free(g->chi);
  free(g);
}

/* visit every element: code = mixed radix (map 0 most significant) */
typedef struct {
  uint64_t best;
  int code, cnt;
  double nre;
  uint64_t s0;
  int mode; /* 0 rep search, 1 check (early exit), 2 norm */
  int rejected;
} qo_vis;

static void qo_visit(const qo_group* g, int k, uint64_t t, int code, qo_vis* v) {
  if (v->rejected) return;
  if (k == g->nmaps) {
    if (v->mode == 0) {
      if (t > v->best) { v->best = t; v->code = code; v->cnt = 1; }
      else if (t == v->best) v->cnt++;
    } else {
      if (v->mode == 1 && t > v->s0) { v->rejected = 1; return; }
      if (t == v->s0) v->nre += g->chr[code];
    }
    return;
  }
  for (int l = 0; l < g->per[k]; ++l) {
    qo_visit(g, k + 1, t, code * g->per[k] + l, v);
    t = qo_apply(g, k, t);
  }
}

uint64_t qo_representative(const qo_group* g, uint64_t s, int* code, int* cnt) {
  qo_vis v = {s, 0, 0, 0.0, s, 0, 0};
  qo_visit(g, 0, s, 0, &v);
  if (code) *code = v.code;
  if (cnt) *cnt = v.cnt;
  return v.best;
}

int qo_normalization(const qo_group* g, uint64_t s) {
  qo_vis v = {s, 0, 0, 0.0, s, 2, 0};
  qo_visit(g, 0, s, 0, &v);
  int n = (int)llround(v.nre);
  return n > 0 ? n : 0;
}

static int qo_check(const qo_group* g, uint64_t s) {
  qo_vis v = {s, 0, 0, 0.0, s, 1, 0};
  qo_visit(g, 0, s, 0, &v);
  if (v.rejected) return 0;
  int n = (int)llround(v.nre);
  return n > 0 ? n : 0;
}

static uint64_t next_comb(uint64_t s) {
  uint64_t c = s & (~s + 1), r = s + c;
  return (((r ^ s) >> 2) / c) | r;
}

static double binom_d(int n, int k) {
  if (k < 0 || k > n) return 0;
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static uint64_t unrank(int64_t r, int k, int N) {
  uint64_t s = 0;
  for (int j = k; j >= 1; --j) {
    int p = j - 1;
    while (p + 1 < N && (int64_t)llround(binom_d(p + 1, j)) <= r) ++p;
    s |= 1ull << p;
    r -= (int64_t)llround(binom_d(p, j));
  }
  return s;
}

/* Representatives of the Np-particle sector, descending, with n.
 * Returns Ns, or -1 if cap is too small. */
int64_t qo_make_basis(const qo_group* g, int Np, uint64_t* basis, int32_t* nout, int64_t cap) {
  const int N = g->N;
  const int64_t total = (int64_t)llround(binom_d(N, Np));
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  const int64_t nchunk = nth * 16;
  int64_t* cnt = (int64_t*)calloc(nchunk + 1, sizeof(int64_t));
  uint64_t** bufs = (uint64_t**)calloc(nchunk, sizeof(uint64_t*));
  int32_t** nbufs = (int32_t**)calloc(nchunk, sizeof(int32_t*));
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t c = 0; c < nchunk; ++c) {
    const int64_t a = total * c / nchunk, b = total * (c + 1) / nchunk;
    if (b <= a) continue;
    int64_t capc = 1024, k = 0;
    uint64_t* buf = (uint64_t*)malloc(capc * sizeof(uint64_t));
    int32_t* nb = (int32_t*)malloc(capc * sizeof(int32_t));
    uint64_t s = Np == 0 ? 0 : unrank(a, Np, N);
    for (int64_t i = a; i < b; ++i) {
      int n = g->nmaps ? qo_check(g, s) : 1;
      if (n > 0) {
        if (k == capc) {
          capc *= 2;
          buf = (uint64_t*)realloc(buf, capc * sizeof(uint64_t));
          nb = (int32_t*)realloc(nb, capc * sizeof(int32_t));
        }
        buf[k] = s;
        nb[k] = n;
        ++k;
      }
      if (i + 1 < b && Np > 0) s = next_comb(s);
    }
    bufs[c] = buf;
    nbufs[c] = nb;
    cnt[c] = k;
  }
  int64_t Ns = 0;
  for (int64_t c = 0; c < nchunk; ++c) Ns += cnt[c];
  int64_t ret = Ns;
  if (basis && Ns <= cap) {
    int64_t pos = Ns - 1;
    for (int64_t c = 0; c < nchunk; ++c)
      for (int64_t k = 0; k < cnt[c]; ++k) {
        basis[pos] = bufs[c][k];
        nout[pos] = nbufs[c][k];
        --pos;
      }
  } else if (basis) {
    ret = -1;
  }
  for (int64_t c = 0; c < nchunk; ++c) { free(bufs[c]); free(nbufs[c]); }
  free(bufs);
  free(nbufs);
  free(cnt);
  return ret;
}

static int64_t qo_index(const uint64_t* b, int64_t Ns, uint64_t key) {
  int64_t lo = 0, hi = Ns;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (b[mid] == key) return mid;
    if (b[mid] > key) lo = mid + 1; else hi = mid;
  }
  return -1;
}

/* hcb spin-1/2 term: ops in opstr order, applied right to left.
 * codes: 'I' 'z' 'n' '+' '-' 'x' ; returns 0 when the amplitude vanishes. */
static int qo_apply_term(int N, const char* ops, const int* sites, int nops, uint64_t* s, double* m) {
  double a = 1.0;
  uint64_t t = *s;
  for (int k = nops - 1; k >= 0; --k) {
    const int p = N - 1 - sites[k];
    const int b = (int)((t >> p) & 1ull);
    switch (ops[k]) {
      case 'I': break;
      case 'z': a *= b ? 0.5 : -0.5; break;
      case 'n': if (!b) return 0; break;
      case '+': if (b) return 0; t ^= 1ull << p; break;
      case '-': if (!b) return 0; t ^= 1ull << p; break;
      case 'x': a *= 0.5; t ^= 1ull << p; break;
      default: return 0;
    }
  }
  *s = t;
  *m = a;
  return 1;
}

/* Real H.v for hcb spin-1/2 (real characters), v_out = H v_in.
 * Terms: nterms strings of length lens[t] in ops (concatenated), sites
 * likewise, couplings J[t].  fused = 1: all terms per row in one pass;
 * fused = 0: one pass over the rows per term (the reference's loop over
 * terms).  Rows gather with the Hermitian-conjugate term.  */
void qo_hv_real(const qo_group* g, const uint64_t* basis, const int32_t* nrm, int64_t Ns, int nterms,
                const char* ops, const int32_t* lens, const int32_t* sites, const double* J, const double* vin,
                double* vout, int fused) {
  const int N = g->N;
  int off[4097];
  off[0] = 0;
  for (int t = 0; t < nterms; ++t) off[t + 1] = off[t] + lens[t];
  /* Hermitian conjugates: reversed, + <-> - */
  char* hops = (char*)malloc(off[nterms] + 1);
  int* hsites = (int*)malloc(sizeof(int) * (off[nterms] + 1));
  for (int t = 0; t < nterms; ++t)
    for (int k = 0; k < lens[t]; ++k) {
      char c = ops[off[t] + lens[t] - 1 - k];
      hops[off[t] + k] = c == '+' ? '-' : (c == '-' ? '+' : c);
      hsites[off[t] + k] = sites[off[t] + lens[t] - 1 - k];
    }
  if (fused) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < Ns; ++r) {
      const uint64_t s = basis[r];
      const double nr = g->nmaps ? (double)nrm[r] : 1.0;
      double acc = 0.0;
      for (int t = 0; t < nterms; ++t) {
        uint64_t s2 = s;
        double m;
        if (!qo_apply_term(N, hops + off[t], hsites + off[t], lens[t], &s2, &m)) continue;
        int code = 0, cnt = 1;
        uint64_t rep = g->nmaps ? qo_representative(g, s2, &code, &cnt) : s2;
        int64_t c = qo_index(basis, Ns, rep);
        if (c < 0) continue;
        double w = J[t] * m * (g->nmaps ? g->chr[code] * sqrt((double)cnt / nr) : 1.0);
        acc += w * vin[c];
      }
      vout[r] = acc;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < Ns; ++r) vout[r] = 0.0;
    for (int t = 0; t < nterms; ++t) {
#pragma omp parallel for schedule(dynamic, 256)
      for (int64_t r = 0; r < Ns; ++r) {
        const uint64_t s = basis[r];
        uint64_t s2 = s;
        double m;
        if (!qo_apply_term(N, hops + off[t], hsites + off[t], lens[t], &s2, &m)) continue;
        int code = 0, cnt = 1;
        uint64_t rep = g->nmaps ? qo_representative(g, s2, &code, &cnt) : s2;
        int64_t c = qo_index(basis, Ns, rep);
        if (c < 0) continue;
        const double nr = g->nmaps ? (double)nrm[r] : 1.0;
        vout[r] += J[t] * m * (g->nmaps ? g->chr[code] * sqrt((double)cnt / nr) : 1.0) * vin[c];
      }
    }
  }
  free(hops);
  free(hsites);
}

/* complex128 CSR SpMV y = A x (row-parallel). */
void qo_csr_spmv_c128(int64_t Ns, const int64_t* indptr, const int64_t* indices, const double* data_ri,
                      const double* x_ri, double* y_ri) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < Ns; ++r) {
    double sr = 0, si = 0;
    for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
      const int64_t c = indices[k];
      const double ar = data_ri[2 * k], ai = data_ri[2 * k + 1];
      const double xr = x_ri[2 * c], xi = x_ri[2 * c + 1];
      sr += ar * xr - ai * xi;
      si += ar * xi + ai * xr;
    }
    y_ri[2 * r] = sr;
    y_ri[2 * r + 1] = si;
  }
}

/* Taylor-series expm on complex128 CSR (scipy _expm_multiply.py:280-303 with
 * A -> A - mu I): F <- e^{a A} F in place; work holds 2*Ns complex. */
void qo_expm_c128(int64_t Ns, const int64_t* indptr, const int64_t* indices, const double* data_ri, int s,
                  int m_star, double tol, double mu_r, double mu_i, double a_r, double a_i, double* F,
                  double* work) {
  double* B = work;
  double* B2 = work + 2 * Ns;
  const double er = a_r * mu_r - a_i * mu_i, ei = a_r * mu_i + a_i * mu_r;
  const double mag = exp(er / s);
  const double eta_r = mag * cos(ei / s), eta_i = mag * sin(ei / s);
  memcpy(B, F, sizeof(double) * 2 * Ns);
  for (int i = 0; i < s; ++i) {
    double c1 = 0;
#pragma omp parallel for reduction(max : c1)
    for (int64_t r = 0; r < Ns; ++r) c1 = fmax(c1, hypot(B[2 * r], B[2 * r + 1]));
    for (int j = 0; j < m_star; ++j) {
      const double cr = a_r / (s * (j + 1.0)), ci = a_i / (s * (j + 1.0));
      qo_csr_spmv_c128(Ns, indptr, indices, data_ri, B, B2);
      double c2 = 0, nf = 0;
#pragma omp parallel for reduction(max : c2, nf)
      for (int64_t r = 0; r < Ns; ++r) {
        const double yr = B2[2 * r] - (mu_r * B[2 * r] - mu_i * B[2 * r + 1]);
        const double yi = B2[2 * r + 1] - (mu_r * B[2 * r + 1] + mu_i * B[2 * r]);
        const double br = cr * yr - ci * yi, bi = cr * yi + ci * yr;
        B2[2 * r] = br;
        B2[2 * r + 1] = bi;
        F[2 * r] += br;
        F[2 * r + 1] += bi;
        c2 = fmax(c2, hypot(br, bi));
        nf = fmax(nf, hypot(F[2 * r], F[2 * r + 1]));
      }
      double* t = B; B = B2; B2 = t;
      if (c1 + c2 <= tol * nf) break;
      c1 = c2;
    }
#pragma omp parallel for
    for (int64_t r = 0; r < Ns; ++r) {
      const double fr = F[2 * r], fi = F[2 * r + 1];
      F[2 * r] = eta_r * fr - eta_i * fi;
      F[2 * r + 1] = eta_r * fi + eta_i * fr;
    }
    memcpy(B, F, sizeof(double) * 2 * Ns);
  }
}

/* Sector-table form of the fused H.v (the lookup the device kernels use,
 * restated on the CPU so that the timed CPU path runs the same algorithm):
 * for every state s of the Np-particle sector, indexed by its colex rank
 * sum_j C(p_j, j + 1) over the set-bit positions p_0 < p_1 < ..., the table
 * holds the basis index of rep(s) (-1 when absent), the group element code
 * reaching it and the number of elements that do (the oracle's cnt).
 * Built with qo_representative; 8 bytes per entry. */
typedef struct {
  int32_t idx;
  uint16_t code, cnt;
} qo_tab_entry;

static int64_t qo_binom_i[65][65];

static void qo_binom_init(void) {
  if (qo_binom_i[1][0]) return;
  for (int n = 0; n <= 64; ++n)
    for (int k = 0; k <= 64; ++k) qo_binom_i[n][k] = k > n ? 0 : (k == 0 || k == n ? 1 : qo_binom_i[n - 1][k - 1] + qo_binom_i[n - 1][k]);
}

static int64_t qo_colex_rank(uint64_t s) {
  int64_t r = 0;
  int j = 1;
  while (s) {
    const int p = __builtin_ctzll(s);
    s &= s - 1;
    r += qo_binom_i[p][j];
    ++j;
  }
  return r;
}

/* Fills tab (C(N, Np) entries); returns the entry count, -1 on bad sizes. */
int64_t qo_table_build(const qo_group* g, int Np, const uint64_t* basis, int64_t Ns, void* tab_out) {
  qo_binom_init();
  const int N = g->N;
  if (N > 62 || Ns >= INT32_MAX || g->G > 65535) return -1;
  qo_tab_entry* tab = (qo_tab_entry*)tab_out;
  const int64_t total = qo_binom_i[N][Np];
  const int64_t chunk = 4096;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t a = 0; a < total; a += chunk) {
    const int64_t b = a + chunk < total ? a + chunk : total;
    uint64_t s = Np == 0 ? 0 : unrank(a, Np, N);
    for (int64_t r = a; r < b; ++r) {
      int code = 0, cnt = 1;
      const uint64_t rep = qo_representative(g, s, &code, &cnt);
      const int64_t c = qo_index(basis, Ns, rep);
      tab[r].idx = (int32_t)c;
      tab[r].code = (uint16_t)code;
      tab[r].cnt = (uint16_t)cnt;
      if (r + 1 < b && Np > 0) s = next_comb(s);
    }
  }
  return total;
}

/* Real fused H.v through the sector table: per row and term the image s2,
 * its rank, the table entry, and J m chi(code) sqrt(cnt / n_r) v[idx]. */
void qo_hv_table_real(const qo_group* g, const uint64_t* basis, const int32_t* nrm, int64_t Ns, const void* tab_in,
                      int nterms, const char* ops, const int32_t* lens, const int32_t* sites, const double* J,
                      const double* vin, double* vout) {
  qo_binom_init();
  const qo_tab_entry* tab = (const qo_tab_entry*)tab_in;
  const int N = g->N;
  int off[4097];
  off[0] = 0;
  for (int t = 0; t < nterms; ++t) off[t + 1] = off[t] + lens[t];
  char* hops = (char*)malloc(off[nterms] + 1);
  int* hsites = (int*)malloc(sizeof(int) * (off[nterms] + 1));
  for (int t = 0; t < nterms; ++t)
    for (int k = 0; k < lens[t]; ++k) {
      char c = ops[off[t] + lens[t] - 1 - k];
      hops[off[t] + k] = c == '+' ? '-' : (c == '-' ? '+' : c);
      hsites[off[t] + k] = sites[off[t] + lens[t] - 1 - k];
    }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < Ns; ++r) {
    const uint64_t s = basis[r];
    const double rn = 1.0 / sqrt((double)nrm[r]);
    double acc = 0.0;
    for (int t = 0; t < nterms; ++t) {
      uint64_t s2 = s;
      double m;
      if (!qo_apply_term(N, hops + off[t], hsites + off[t], lens[t], &s2, &m)) continue;
      const qo_tab_entry e = tab[qo_colex_rank(s2)];
      if (e.idx < 0) continue;
      acc += J[t] * m * g->chr[e.code] * sqrt((double)e.cnt) * rn * vin[e.idx];
    }
    vout[r] = acc;
  }
  free(hops);
  free(hsites);
}

/* Real-valued CSR (int32 or int64 index arrays, idx64 = 1 for int64) plus
 * an optional diagonal dscale * D, applied to complex128 vectors: the form
 * in which the reference's OpenMP Taylor loop runs a real Hamiltonian on a
 * complex state (EMP.html:347-372; scipy _expm_multiply.py:280-303). */
typedef struct {
  int64_t Ns;
  const void* indptr;
  const void* indices;
  int idx64;
  const double* data;
  const double* D;
  double dscale;
} qo_rcsr;

static inline int64_t rc_ip(const qo_rcsr* A, int64_t r) {
  return A->idx64 ? ((const int64_t*)A->indptr)[r] : (int64_t)((const int32_t*)A->indptr)[r];
}

static inline int64_t rc_ix(const qo_rcsr* A, int64_t k) {
  return A->idx64 ? ((const int64_t*)A->indices)[k] : (int64_t)((const int32_t*)A->indices)[k];
}

static void rcsr_spmv(const qo_rcsr* A, const double* x, double* y) {
#pragma omp parallel for schedule(static, 4096)
  for (int64_t r = 0; r < A->Ns; ++r) {
    double sr = 0, si = 0;
    const int64_t e = rc_ip(A, r + 1);
    for (int64_t k = rc_ip(A, r); k < e; ++k) {
      const int64_t c = rc_ix(A, k);
      const double a = A->data[k];
      sr += a * x[2 * c];
      si += a * x[2 * c + 1];
    }
    if (A->D) {
      const double d = A->dscale * A->D[r];
      sr += d * x[2 * r];
      si += d * x[2 * r + 1];
    }
    y[2 * r] = sr;
    y[2 * r + 1] = si;
  }
}

/* max_j sum_i |A_ij + delta_ij (dscale D_j - mu)| (the reference's
 * _wrapper_csr_1_norm of A - mu I, EMP.html:200-202), column sums by atomic
 * accumulation. colsum: Ns doubles of scratch. */
double qo_rcsr_1_norm(int64_t Ns, const void* indptr, const void* indices, int idx64, const double* data,
                      const double* D, double dscale, double mu_r, double mu_i, double* colsum) {
  qo_rcsr A = {Ns, indptr, indices, idx64, data, D, dscale};
  double* dg = (double*)calloc(Ns > 0 ? Ns : 1, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < Ns; ++r) {
    colsum[r] = 0.0;
    dg[r] = D ? dscale * D[r] : 0.0;
  }
#pragma omp parallel for schedule(static, 4096)
  for (int64_t r = 0; r < Ns; ++r) {
    const int64_t e = rc_ip(&A, r + 1);
    for (int64_t k = rc_ip(&A, r); k < e; ++k) {
      const int64_t c = rc_ix(&A, k);
      if (c == r) {
        dg[r] += data[k];
      } else {
        const double v = fabs(data[k]);
#pragma omp atomic
        colsum[c] += v;
      }
    }
  }
  double best = 0.0;
#pragma omp parallel for reduction(max : best)
  for (int64_t j = 0; j < Ns; ++j) {
    const double v = colsum[j] + hypot(dg[j] - mu_r, -mu_i);
    best = fmax(best, v);
  }
  free(dg);
  return best;
}

/* Taylor loop of qo_expm_c128 on the real CSR form (plus diagonal). */
void qo_expm_rcsr_c128(int64_t Ns, const void* indptr, const void* indices, int idx64, const double* data,
                       const double* D, double dscale, int s, int m_star, double tol, double mu_r, double mu_i,
                       double a_r, double a_i, double* F, double* work) {
  qo_rcsr A = {Ns, indptr, indices, idx64, data, D, dscale};
  double* B = work;
  double* B2 = work + 2 * Ns;
  const double er = a_r * mu_r - a_i * mu_i, ei = a_r * mu_i + a_i * mu_r;
  const double mag = exp(er / s);
  const double eta_r = mag * cos(ei / s), eta_i = mag * sin(ei / s);
  memcpy(B, F, sizeof(double) * 2 * Ns);
  for (int i = 0; i < s; ++i) {
    double c1 = 0;
#pragma omp parallel for reduction(max : c1)
    for (int64_t r = 0; r < Ns; ++r) c1 = fmax(c1, hypot(B[2 * r], B[2 * r + 1]));
    for (int j = 0; j < m_star; ++j) {
      const double cr = a_r / (s * (j + 1.0)), ci = a_i / (s * (j + 1.0));
      rcsr_spmv(&A, B, B2);
      double c2 = 0, nf = 0;
#pragma omp parallel for reduction(max : c2, nf)
      for (int64_t r = 0; r < Ns; ++r) {
        const double yr = B2[2 * r] - (mu_r * B[2 * r] - mu_i * B[2 * r + 1]);
        const double yi = B2[2 * r + 1] - (mu_r * B[2 * r + 1] + mu_i * B[2 * r]);
        const double br = cr * yr - ci * yi, bi = cr * yi + ci * yr;
        B2[2 * r] = br;
        B2[2 * r + 1] = bi;
        F[2 * r] += br;
        F[2 * r + 1] += bi;
        c2 = fmax(c2, hypot(br, bi));
        nf = fmax(nf, hypot(F[2 * r], F[2 * r + 1]));
      }
      double* t = B; B = B2; B2 = t;
      if (c1 + c2 <= tol * nf) break;
      c1 = c2;
    }
#pragma omp parallel for
    for (int64_t r = 0; r < Ns; ++r) {
      const double fr = F[2 * r], fi = F[2 * r + 1];
      F[2 * r] = eta_r * fr - eta_i * fi;
      F[2 * r + 1] = eta_r * fi + eta_i * fr;
    }
    memcpy(B, F, sizeof(double) * 2 * Ns);
  }
}

int qo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
