// Batched sectors x vectors: one launch applies H_k to the stacked vectors of
// every sector k, and the Taylor loop of e^{a H_k} V_k runs for all sectors in
// lockstep (one launch per Taylor term for the whole batch, per-sector
// partition, truncation test and done flag), optionally replayed from a
// captured hipGraph.  Extends the per-sector calls
// ``expm_multiply_parallel(H_k, a).dot(V_k)`` (EMP.html:267-374 with
// ``_wrapper_expm_multiply_batch``) to many sectors at once; the result of
// every sector equals its own call.
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "batch_impl.h"

struct qsx_oplist_full;  // defined in hv.hip
struct qsx_hv_ctx;
int qsx_hv_ctx_export(const qsx_hv_ctx* X, HvArgs* A, int* dtype, int* N, int* lookup, int* fermion);

struct qsx_hv_batch {
  int nsec = 0, dtype = 0, N = 0, lookup = 0;
  bool fermion = false;
  int64_t nvec = 1, total = 0;
  int chunk_rows = 16;
  int xcd_slices = 0;         // 1: each XCD takes a contiguous slice of the chunks
  std::vector<int64_t> row_off_h;
  HvArgs* sec_d = nullptr;
  CsrSec* csr_d = nullptr;    // CSR form
  int64_t* row_off_d = nullptr;
  WaveChunk* chunk_d = nullptr;
  int64_t nchunk = 0;
  // Taylor-loop state
  ExpmSec* es_d = nullptr;
  std::vector<ExpmSec> es_h;  // host copy kept alive for the asynchronous upload
  int* done_d = nullptr;
  double* slots_d = nullptr;
  int64_t slots_cap = 0;
  // captured Taylor loops keyed by (v, work, s_max, m_max)
  hipStream_t cap = nullptr;
  std::map<std::tuple<const void*, const void*, int, int>, hipGraphExec_t> graphs;
};

// zero every norm slot and done flag
__global__ void k_batch_init(double* nB, double* nF, int64_t nslots, int* done, int nsec) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (int64_t)gridDim.x * blockDim.x) {
    nB[i] = 0.0;
    nF[i] = 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsec; i += (int64_t)gridDim.x * blockDim.x)
    done[i] = 0;
}

static void free_batch(qsx_hv_batch* b) {
  if (!b) return;
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  (void)hipFree(b->sec_d);
  (void)hipFree(b->csr_d);
  (void)hipFree(b->row_off_d);
  (void)hipFree(b->chunk_d);
  (void)hipFree(b->es_d);
  (void)hipFree(b->done_d);
  (void)hipFree(b->slots_d);
  delete b;
}

static int build_chunks(qsx_hv_batch* b) {
  std::vector<WaveChunk> ch;
  for (int k = 0; k < b->nsec; ++k) {
    const int64_t ns = b->row_off_h[k + 1] - b->row_off_h[k];
    for (int64_t r = 0; r < ns; r += b->chunk_rows) {
      WaveChunk c;
      c.sec = k;
      c.n = (int32_t)std::min<int64_t>(b->chunk_rows, ns - r);
      c.begin = r;
      ch.push_back(c);
    }
  }
  (void)hipFree(b->chunk_d);
  b->chunk_d = nullptr;
  b->nchunk = (int64_t)ch.size();
  if (ch.empty()) return QSX_OK;
  QSX_CHECK_HIP(hipMalloc((void**)&b->chunk_d, sizeof(WaveChunk) * ch.size()));
  QSX_CHECK_HIP(hipMemcpy(b->chunk_d, ch.data(), sizeof(WaveChunk) * ch.size(), hipMemcpyHostToDevice));
  // graphs hold the old chunk pointer
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  return QSX_OK;
}

extern "C" int qsx_hv_batch_create(int nsec, qsx_hv_ctx* const* ctxs, int64_t nvec, int chunk_rows,
                                   qsx_hv_batch** out) {
  if (nsec < 1 || !ctxs || !out || nvec < 1 || chunk_rows < 1 || chunk_rows > (1 << 20)) {
    qsx_set_error("qsx_hv_batch_create: bad arguments");
    return QSX_ERR_ARG;
  }
  auto* b = new qsx_hv_batch();
  b->nsec = nsec;
  b->nvec = nvec;
  b->chunk_rows = chunk_rows;
  std::vector<HvArgs> sec(nsec);
  b->row_off_h.assign(nsec + 1, 0);
  for (int k = 0; k < nsec; ++k) {
    int dt, N, lk, fm;
    int rc = qsx_hv_ctx_export(ctxs[k], &sec[k], &dt, &N, &lk, &fm);
    if (rc) { free_batch(b); return rc; }
    if (k == 0) { b->dtype = dt; b->N = N; b->lookup = lk; b->fermion = fm != 0; }
    if (dt != b->dtype || N != b->N || lk != b->lookup || (fm != 0) != b->fermion) {
      qsx_set_error("qsx_hv_batch_create: sectors differ in dtype, N, lookup or statistics");
      free_batch(b);
      return QSX_ERR_ARG;
    }
    if (sec[k].Ns >= ((int64_t)1 << 31)) {
      qsx_set_error("qsx_hv_batch_create: sectors of 2^31 states or more are not batched");
      free_batch(b);
      return QSX_ERR_UNSUPPORTED;
    }
    b->row_off_h[k + 1] = b->row_off_h[k] + sec[k].Ns;
  }
  b->total = b->row_off_h[nsec];
  hipError_t e = hipMalloc((void**)&b->sec_d, sizeof(HvArgs) * nsec);
  if (e == hipSuccess) e = hipMemcpy(b->sec_d, sec.data(), sizeof(HvArgs) * nsec, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&b->row_off_d, sizeof(int64_t) * (nsec + 1));
  if (e == hipSuccess)
    e = hipMemcpy(b->row_off_d, b->row_off_h.data(), sizeof(int64_t) * (nsec + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&b->es_d, sizeof(ExpmSec) * nsec);
  if (e == hipSuccess) e = hipMalloc((void**)&b->done_d, sizeof(int) * nsec);
  if (e != hipSuccess) {
    qsx_set_error(std::string("qsx_hv_batch_create: ") + hipGetErrorString(e));
    free_batch(b);
    return QSX_ERR_HIP;
  }
  int rc = build_chunks(b);
  if (rc) { free_batch(b); return rc; }
  *out = b;
  return QSX_OK;
}

// CSR form of a batch: sector k is the CSR matrix (indptr[k] with Ns[k]+1
// entries, int64 when idx64[k], int32 indices[k], data[k] of the vector
// dtype), all on the device and owned by the caller.
extern "C" int qsx_csr_batch_create(int nsec, const int64_t* Ns, const void* const* indptr, const int32_t* idx64,
                                    const void* const* indices, const void* const* data, int dtype, int64_t nvec,
                                    int chunk_rows, qsx_hv_batch** out) {
  if (nsec < 1 || !Ns || !indptr || !idx64 || !indices || !data || !out || nvec < 1 || chunk_rows < 1 ||
      chunk_rows > 63 || dtype < QSX_F32 || dtype > QSX_C128) {
    qsx_set_error("qsx_csr_batch_create: bad arguments (chunk_rows must be 1..63)");
    return QSX_ERR_ARG;
  }
  auto* b = new qsx_hv_batch();
  b->nsec = nsec;
  b->nvec = nvec;
  b->chunk_rows = chunk_rows;
  b->dtype = dtype;
  std::vector<CsrSec> sec(nsec);
  b->row_off_h.assign(nsec + 1, 0);
  for (int k = 0; k < nsec; ++k) {
    if (Ns[k] < 0 || Ns[k] >= ((int64_t)1 << 31) || (Ns[k] > 0 && (!indptr[k] || !indices[k] || !data[k]))) {
      qsx_set_error("qsx_csr_batch_create: bad sector matrix");
      free_batch(b);
      return QSX_ERR_ARG;
    }
    std::memset(&sec[k], 0, sizeof(CsrSec));
    sec[k].indptr = indptr[k];
    sec[k].indices = (const int32_t*)indices[k];
    sec[k].data = data[k];
    sec[k].Ns = Ns[k];
    sec[k].idx64 = idx64[k] ? 1 : 0;
    b->row_off_h[k + 1] = b->row_off_h[k] + Ns[k];
  }
  b->total = b->row_off_h[nsec];
  hipError_t e = hipMalloc((void**)&b->csr_d, sizeof(CsrSec) * nsec);
  if (e == hipSuccess) e = hipMemcpy(b->csr_d, sec.data(), sizeof(CsrSec) * nsec, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&b->row_off_d, sizeof(int64_t) * (nsec + 1));
  if (e == hipSuccess)
    e = hipMemcpy(b->row_off_d, b->row_off_h.data(), sizeof(int64_t) * (nsec + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&b->es_d, sizeof(ExpmSec) * nsec);
  if (e == hipSuccess) e = hipMalloc((void**)&b->done_d, sizeof(int) * nsec);
  if (e != hipSuccess) {
    qsx_set_error(std::string("qsx_csr_batch_create: ") + hipGetErrorString(e));
    free_batch(b);
    return QSX_ERR_HIP;
  }
  int rc = build_chunks(b);
  if (rc) { free_batch(b); return rc; }
  *out = b;
  return QSX_OK;
}

extern "C" void qsx_hv_batch_destroy(qsx_hv_batch* b) { free_batch(b); }

extern "C" int qsx_hv_batch_rows(const qsx_hv_batch* b, int64_t* row_off) {
  if (!b || !row_off) { qsx_set_error("qsx_hv_batch_rows: bad arguments"); return QSX_ERR_ARG; }
  for (int k = 0; k <= b->nsec; ++k) row_off[k] = b->row_off_h[k];
  return b->nsec;
}

// Rows per wavefront chunk and chunk order (tuning knobs; results do not
// depend on them).
extern "C" int qsx_hv_batch_set_schedule(qsx_hv_batch* b, int chunk_rows, int xcd_slices) {
  if (!b || chunk_rows < 1 || chunk_rows > (1 << 20) || (b->csr_d && chunk_rows > 63)) {
    qsx_set_error("qsx_hv_batch_set_schedule: bad arguments (CSR batches take at most 63 rows per chunk)");
    return QSX_ERR_ARG;
  }
  if (xcd_slices != b->xcd_slices) {
    b->xcd_slices = xcd_slices ? 1 : 0;
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
  }
  if (chunk_rows == b->chunk_rows) return QSX_OK;
  b->chunk_rows = chunk_rows;
  return build_chunks(b);
}

static dim3 wave_grid(const qsx_hv_batch* b, int* xcd_blocks) {
  const int64_t blocks = (b->nchunk + 3) / 4;
  if (!b->xcd_slices) {
    *xcd_blocks = 0;
    return dim3((unsigned)blocks);
  }
  const int64_t per = (blocks + 7) / 8;
  *xcd_blocks = (int)per;
  return dim3((unsigned)(per * 8));
}

static BatchArgs base_args(const qsx_hv_batch* b) {
  BatchArgs B;
  std::memset(&B, 0, sizeof(B));
  B.sec = b->sec_d;
  B.csr = b->csr_d;
  B.row_off = b->row_off_d;
  B.chunk = b->chunk_d;
  B.nchunk = b->nchunk;
  B.ld = b->nvec;
  B.xdiag = nullptr;
  B.xdscale = make_double2(0.0, 0.0);
  return B;
}

static void wave_launch(const qsx_hv_batch* b, const BatchArgs& B, int epi, int stage_kind, dim3 grid,
                        hipStream_t st) {
  switch (b->dtype) {
    case QSX_F32: qsx_wave_launch_f32(B, b->N, b->lookup, b->fermion, epi, stage_kind, grid, st); break;
    case QSX_F64: qsx_wave_launch_f64(B, b->N, b->lookup, b->fermion, epi, stage_kind, grid, st); break;
    case QSX_C64: qsx_wave_launch_c64(B, b->N, b->lookup, b->fermion, epi, stage_kind, grid, st); break;
    default: qsx_wave_launch_c128(B, b->N, b->lookup, b->fermion, epi, stage_kind, grid, st); break;
  }
}

// y = alpha * H x (+ y if beta), x and y stacked (total_rows, nvec) on the device.
extern "C" int qsx_hv_batch_apply(qsx_hv_batch* b, const void* x, void* y, double a_re, double a_im, int beta,
                                  void* stream) {
  if (!b || !x || !y) { qsx_set_error("qsx_hv_batch_apply: bad arguments"); return QSX_ERR_ARG; }
  if (b->nchunk == 0) return QSX_OK;
  const bool real = (b->dtype == QSX_F32 || b->dtype == QSX_F64);
  if (real && a_im != 0.0) { qsx_set_error("real vectors with a complex scale"); return QSX_ERR_COMPLEX; }
  hipStream_t st = (hipStream_t)stream;
  int xb;
  const dim3 grid = wave_grid(b, &xb);
  for (int64_t c0 = 0; c0 < b->nvec; c0 += 64) {
    BatchArgs B = base_args(b);
    B.xcd_blocks = xb;
    B.col0 = c0;
    B.ncol = (int)std::min<int64_t>(64, b->nvec - c0);
    B.x = x;
    B.y = y;
    B.alpha = make_double2(a_re, a_im);
    B.beta = beta ? 1 : 0;
    wave_launch(b, B, 0, 0, grid, st);
    QSX_LAUNCH_CHECK();
  }
  return QSX_OK;
}

// Per-sector partitions s[k], m[k], tolerances tol[k], shifts mu[2k], a common
// a; v = stacked F (overwritten with e^{a H_k} V_k), work = 2 * total * nvec.
// use_graph = 1 records the launch sequence once per (v, work, s_max, m_max)
// and replays it; the per-sector parameters live in device memory, so a
// replay picks up new partitions of the same maximal shape.
extern "C" int qsx_expm_multiply_batch(qsx_hv_batch* b, const int32_t* s, const int32_t* m, const double* tol,
                                       const double* mu, double a_re, double a_im, void* v, void* work,
                                       int use_graph, void* stream) {
  if (!b || !s || !m || !tol || !mu || !v || !work) { qsx_set_error("qsx_expm_multiply_batch: bad arguments"); return QSX_ERR_ARG; }
  if (b->nvec > 64) { qsx_set_error("qsx_expm_multiply_batch: at most 64 vectors per sector"); return QSX_ERR_UNSUPPORTED; }
  const bool real = (b->dtype == QSX_F32 || b->dtype == QSX_F64);
  hipStream_t st = (hipStream_t)stream;
  int smax = 1, mmax = 0;
  if (b->nchunk > 0) QSX_CHECK_HIP(hipStreamSynchronize(st));  // the previous upload has been consumed
  std::vector<ExpmSec>& es = b->es_h;
  es.assign(b->nsec, ExpmSec{});
  for (int k = 0; k < b->nsec; ++k) {
    if (s[k] < 1 || m[k] < 0) { qsx_set_error("qsx_expm_multiply_batch: bad partition"); return QSX_ERR_ARG; }
    if (real && (a_im != 0.0 || mu[2 * k + 1] != 0.0)) { qsx_set_error("real vectors with complex a or mu"); return QSX_ERR_COMPLEX; }
    smax = std::max(smax, (int)s[k]);
    mmax = std::max(mmax, (int)m[k]);
    ExpmSec& E = es[k];
    std::memset(&E, 0, sizeof(E));
    E.a = make_double2(a_re, a_im);
    E.mu = make_double2(mu[2 * k], mu[2 * k + 1]);
    E.dscale = make_double2(0.0, 0.0);
    E.tol = tol[k];
    E.s = s[k];
    E.m = m[k];
  }
  if (b->nchunk == 0) return QSX_OK;
  const int64_t nslot = (int64_t)smax * (mmax + 1);
  const int64_t need = 2 * nslot * b->nsec;
  if (need > b->slots_cap) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    QSX_CHECK_HIP(hipStreamSynchronize(st));
    (void)hipFree(b->slots_d);
    b->slots_d = nullptr;
    QSX_CHECK_HIP(hipMalloc((void**)&b->slots_d, sizeof(double) * need));
    b->slots_cap = need;
  }
  QSX_CHECK_HIP(hipMemcpyAsync(b->es_d, es.data(), sizeof(ExpmSec) * b->nsec, hipMemcpyHostToDevice, st));

  size_t esz = (b->dtype == QSX_F32) ? 4 : (b->dtype == QSX_F64 || b->dtype == QSX_C64) ? 8 : 16;
  void* Bb[2] = {work, (char*)work + esz * b->total * b->nvec};
  int xb;
  const dim3 grid = wave_grid(b, &xb);
  BatchArgs B0 = base_args(b);
  B0.xcd_blocks = xb;
  B0.ncol = (int)b->nvec;
  B0.F = v;
  B0.es = b->es_d;
  B0.done = b->done_d;
  B0.nB = b->slots_d;
  B0.nF = b->slots_d + nslot * b->nsec;
  B0.nslot = nslot;
  B0.mstride = mmax + 1;

  auto record = [&](hipStream_t q) -> int {
    hipLaunchKernelGGL(k_batch_init, dim3(64), dim3(256), 0, q, B0.nB, B0.nF, nslot * b->nsec, b->done_d, b->nsec);
    QSX_LAUNCH_CHECK();
    BatchArgs S = B0;
    S.y = Bb[0];
    S.stage = 0;
    wave_launch(b, S, 0, 1, grid, q);
    QSX_LAUNCH_CHECK();
    for (int i = 0; i < smax; ++i) {
      for (int j = 0; j < mmax; ++j) {
        BatchArgs T = B0;
        T.x = Bb[j & 1];
        T.y = Bb[(j + 1) & 1];
        T.stage = i;
        T.term = j;
        wave_launch(b, T, 1, 0, grid, q);
        QSX_LAUNCH_CHECK();
      }
      BatchArgs C = B0;
      C.y = Bb[0];
      C.stage = i;
      wave_launch(b, C, 0, 2, grid, q);
      QSX_LAUNCH_CHECK();
    }
    return QSX_OK;
  };

  if (!use_graph) return record(st);

  const auto key = std::make_tuple((const void*)v, (const void*)work, smax, mmax);
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    if (!b->cap) QSX_CHECK_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    QSX_CHECK_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    int rc = record(b->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { qsx_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return QSX_ERR_HIP; }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { qsx_set_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return QSX_ERR_HIP; }
    it = b->graphs.emplace(key, ex).first;
  }
  QSX_CHECK_HIP(hipGraphLaunch(it->second, st));
  return QSX_OK;
}

// Size in bytes of one ExpmSec record (for callers that fill them).
extern "C" int qsx_expm_record_bytes() { return (int)sizeof(ExpmSec); }
