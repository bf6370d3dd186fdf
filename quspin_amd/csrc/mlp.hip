// Small MLP (4 -> 64 -> 64 -> 2, tanh hidden layers, FP64) that predicts the
// Taylor partition (m, s) of the expm of a time-dependent H(t) from
// (t, dt, drive phase, ||H(t)||_1 bound).  Inference, gradients of the
// mean-squared error and the Adam update are all device kernels; training
// data are partitions chosen by the reference algorithm (scipy
// _expm_multiply.py:503-558 _fragment_3_1 with the exact 1-norm, as
// EMP.html:377-385 calls it).
//
// Parameter layout (row-major): W1[64][4], b1[64], W2[64][64], b2[64],
// W3[2][64], b3[2].  norm[12] = input shift[4], input scale[4], output
// shift[2], output scale[2]: x_hat = (x - shift) / scale, y = y_hat * scale +
// shift.  One wavefront per sample, lane j = hidden unit j.
#include <cmath>

#include "batch_impl.h"

namespace {
constexpr int MI = 4, MH = 64, MO = 2;
constexpr int P_W1 = 0, P_B1 = P_W1 + MH * MI, P_W2 = P_B1 + MH, P_B2 = P_W2 + MH * MH, P_W3 = P_B2 + MH,
              P_B3 = P_W3 + MO * MH, P_TOT = P_B3 + MO;
constexpr int GRAD_WAVES = 256;  // fixed number of per-wave gradient partials (deterministic sums)
}  // namespace

struct MlpFwd {
  double x[MI];   // normalised input (uniform)
  double h1, h2;  // this lane's hidden units
  double y[MO];   // normalised output (uniform)
};

__device__ __forceinline__ MlpFwd mlp_fwd_wave(const double* __restrict__ P, const double* __restrict__ norm,
                                               const double* __restrict__ xin) {
  const int j = lane_id();
  MlpFwd f;
#pragma unroll
  for (int i = 0; i < MI; ++i) f.x[i] = (xin[i] - norm[i]) / norm[MI + i];
  double z1 = P[P_B1 + j];
#pragma unroll
  for (int i = 0; i < MI; ++i) z1 += P[P_W1 + j * MI + i] * f.x[i];
  f.h1 = tanh(z1);
  double z2 = P[P_B2 + j];
  for (int k = 0; k < MH; ++k) z2 += P[P_W2 + j * MH + k] * __shfl(f.h1, k, 64);
  f.h2 = tanh(z2);
#pragma unroll
  for (int o = 0; o < MO; ++o) f.y[o] = wave_sum(P[P_W3 + o * MH + j] * f.h2) + P[P_B3 + o];
  return f;
}

__global__ void __launch_bounds__(256) k_mlp_forward(const double* __restrict__ P, const double* __restrict__ norm,
                                                     const double* __restrict__ X, int64_t n, double* __restrict__ Y) {
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t s = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += nw) {
    const MlpFwd f = mlp_fwd_wave(P, norm, X + s * MI);
    if (lane_id() < MO) {
      const int o = lane_id();
      Y[s * MO + o] = f.y[o] * norm[2 * MI + MO + o] + norm[2 * MI + o];
    }
  }
}

// Per-wave partial gradients of L = 1/(2n) sum_s |y_hat_s - t_hat_s|^2
// (t_hat = normalised target) into part[w][P_TOT + 1] (last slot: loss part).
__global__ void __launch_bounds__(256) k_mlp_grad(const double* __restrict__ P, const double* __restrict__ norm,
                                                  const double* __restrict__ X, const double* __restrict__ T,
                                                  int64_t n, double* __restrict__ part) {
  const int j = lane_id();
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  double gW1[MI] = {0, 0, 0, 0}, gb1 = 0, gW2[MH], gb2 = 0, gW3[MO] = {0, 0}, gb3[MO] = {0, 0}, loss = 0;
#pragma unroll
  for (int k = 0; k < MH; ++k) gW2[k] = 0.0;
  const double inv_n = 1.0 / (double)n;
  for (int64_t s = wid; s < n; s += nw) {
    const MlpFwd f = mlp_fwd_wave(P, norm, X + s * MI);
    double dy[MO];
#pragma unroll
    for (int o = 0; o < MO; ++o) {
      const double t = (T[s * MO + o] - norm[2 * MI + o]) / norm[2 * MI + MO + o];
      const double e = f.y[o] - t;
      loss += 0.5 * e * e * inv_n;
      dy[o] = e * inv_n;
      gW3[o] += dy[o] * f.h2;
      gb3[o] += dy[o];
    }
    double dh2 = 0.0;
#pragma unroll
    for (int o = 0; o < MO; ++o) dh2 += dy[o] * P[P_W3 + o * MH + j];
    dh2 *= 1.0 - f.h2 * f.h2;
#pragma unroll
    for (int k = 0; k < MH; ++k) gW2[k] += dh2 * __shfl(f.h1, k, 64);
    gb2 += dh2;
    // dh1_j = (sum_i dh2_i W2[i][j]) (1 - h1_j^2)
    double dh1 = 0.0;
    for (int i = 0; i < MH; ++i) dh1 += __shfl(dh2, i, 64) * P[P_W2 + i * MH + j];
    dh1 *= 1.0 - f.h1 * f.h1;
#pragma unroll
    for (int i = 0; i < MI; ++i) gW1[i] += dh1 * f.x[i];
    gb1 += dh1;
  }
  if (wid >= GRAD_WAVES) return;
  double* out = part + wid * (P_TOT + 1);
#pragma unroll
  for (int i = 0; i < MI; ++i) out[P_W1 + j * MI + i] = gW1[i];
  out[P_B1 + j] = gb1;
#pragma unroll
  for (int k = 0; k < MH; ++k) out[P_W2 + j * MH + k] = gW2[k];
  out[P_B2 + j] = gb2;
#pragma unroll
  for (int o = 0; o < MO; ++o) out[P_W3 + o * MH + j] = gW3[o];
  if (j == 0) {
#pragma unroll
    for (int o = 0; o < MO; ++o) out[P_B3 + o] = gb3[o];
    out[P_TOT] = loss;
  }
}

// grad[p] = sum over the per-wave partials in wave order (deterministic)
__global__ void k_mlp_reduce(const double* __restrict__ part, int nw, double* __restrict__ grad, double* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P_TOT) return;
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += part[(int64_t)w * (P_TOT + 1) + p];
  if (p < P_TOT) grad[p] = s;
  else if (loss) *loss = s;
}

__global__ void k_mlp_adam(double* __restrict__ P, const double* __restrict__ g, double* __restrict__ m,
                           double* __restrict__ v, double lr, double b1, double b2, double eps, double c1, double c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P_TOT) return;
  const double gp = g[p];
  const double mp = b1 * m[p] + (1.0 - b1) * gp;
  const double vp = b2 * v[p] + (1.0 - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp / c1) / (sqrt(vp / c2) + eps);
}

// Partitions of the time steps: (m, s) from the network, m rounded up and
// padded by m_margin, both clamped; written into the steps' ExpmSec records.
__global__ void __launch_bounds__(256) k_mlp_partition(const double* __restrict__ P, const double* __restrict__ norm,
                                                       const double* __restrict__ X, int64_t n, ExpmSec* es,
                                                       int m_margin, int m_cap, int s_cap) {
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t s = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += nw) {
    const MlpFwd f = mlp_fwd_wave(P, norm, X + s * MI);
    if (lane_id() == 0) {
      const double ym = f.y[0] * norm[2 * MI + MO] + norm[2 * MI];
      const double ys = f.y[1] * norm[2 * MI + MO + 1] + norm[2 * MI + 1];
      int m = (int)ceil(ym) + m_margin, sp = (int)llround(ys);
      m = m < 1 ? 1 : (m > m_cap ? m_cap : m);
      sp = sp < 1 ? 1 : (sp > s_cap ? s_cap : sp);
      es[s].m = m;
      es[s].s = sp;
    }
  }
}

extern "C" int qsx_mlp_param_count() { return P_TOT; }

extern "C" int qsx_mlp_forward(const double* params, const double* norm, const double* X, int64_t n, double* Y,
                               void* stream) {
  if (!params || !norm || (n > 0 && (!X || !Y)) || n < 0) { qsx_set_error("qsx_mlp_forward: bad arguments"); return QSX_ERR_ARG; }
  if (n == 0) return QSX_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 3) / 4, 1024);
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(256), 0, (hipStream_t)stream, params, norm, X, n, Y);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

// work: at least qsx_mlp_grad_work() doubles of device memory.
extern "C" int64_t qsx_mlp_grad_work() { return (int64_t)GRAD_WAVES * (P_TOT + 1); }

extern "C" int qsx_mlp_grad(const double* params, const double* norm, const double* X, const double* T, int64_t n,
                            double* grad, double* loss, double* work, int64_t work_elems, void* stream) {
  if (!params || !norm || !X || !T || !grad || !work || n < 1 || work_elems < qsx_mlp_grad_work()) {
    qsx_set_error("qsx_mlp_grad: bad arguments");
    return QSX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  // GRAD_WAVES waves exactly (4 per block); waves without samples write zeros
  hipLaunchKernelGGL(k_mlp_grad, dim3(GRAD_WAVES / 4), dim3(256), 0, st, params, norm, X, T, n, work);
  QSX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mlp_reduce, dim3((P_TOT + 1 + 255) / 256), dim3(256), 0, st, work, GRAD_WAVES, grad, loss);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_mlp_adam(double* params, const double* grad, double* m, double* v, int64_t step, double lr,
                            double beta1, double beta2, double eps, void* stream) {
  if (!params || !grad || !m || !v || step < 1) { qsx_set_error("qsx_mlp_adam: bad arguments"); return QSX_ERR_ARG; }
  const double c1 = 1.0 - std::pow(beta1, (double)step), c2 = 1.0 - std::pow(beta2, (double)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((P_TOT + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, grad, m, v,
                     lr, beta1, beta2, eps, c1, c2);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}

extern "C" int qsx_mlp_partition(const double* params, const double* norm, const double* X, int64_t n, void* es,
                                 int m_margin, int m_cap, int s_cap, void* stream) {
  if (!params || !norm || !X || !es || n < 1 || m_cap < 1 || s_cap < 1) { qsx_set_error("qsx_mlp_partition: bad arguments"); return QSX_ERR_ARG; }
  const unsigned g = (unsigned)std::min<int64_t>((n + 3) / 4, 1024);
  hipLaunchKernelGGL(k_mlp_partition, dim3(g), dim3(256), 0, (hipStream_t)stream, params, norm, X, n, (ExpmSec*)es,
                     m_margin, m_cap, s_cap);
  QSX_LAUNCH_CHECK();
  return QSX_OK;
}
