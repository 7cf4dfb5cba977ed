// Learned con2prim initial guess for RMHD: training pairs (sampled states or
// the states of a run), FP32 forward/backward training kernels with Adam,
// the matrix-core form of the weights, inference, and a batched Newton solve
// that reports its passes -- behind the C ABI.
#include "../../include/gpluto_amd.h"
#include "c2p_net.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace gpl {

namespace {

// splitmix64 (Steele, Lea & Flood 2014): a counter-based stream per sample
__host__ __device__ __forceinline__ uint64_t smix(uint64_t &s)
{
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27))*0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unif(uint64_t &s) { return (double)(smix(s) >> 11)*0x1.0p-53; }

__device__ __forceinline__ void unit_vector(uint64_t &s, double *d)
{
  double nrm = 0.0;
  for (int tries = 0; tries < 64; tries++) {     // rejection: uniform direction
    for (int q = 0; q < 3; q++) d[q] = 2.0*unif(s) - 1.0;
    nrm = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
    if (nrm <= 1.0 && nrm >= 1e-6) break;
  }
  if (nrm > 1.0 || nrm < 1e-6) { d[0] = 1.0; d[1] = d[2] = 0.0; nrm = 1.0; }
  nrm = sqrt(nrm);
  for (int q = 0; q < 3; q++) d[q] /= nrm;
}

// the training target of a (conserved, primitive) pair: log(W/W_A27), W = rho h gamma^2
__device__ __forceinline__ float net_target(const double *u, const double *v, double gamma)
{
  const double g2 = 1.0/(1.0 - (v[1]*v[1] + v[2]*v[2] + v[3]*v[3]));
  const double h = 1.0 + gamma/(gamma - 1.0)*v[7]/v[0];
  return (float)log(v[0]*h*g2/rmhd_guess_a27(u));
}

// (conserved, primitive) pairs made by the method itself: a primitive state
// drawn over rho in [1e-2,1e2], p/rho in [1e-3,1e2], Lorentz factor in
// [1,20], sigma = B^2/(rho h) in [1e-3,1e2] (log-uniform, isotropic
// directions), mapped by PrimToCons.
__global__ __launch_bounds__(256)
void k_net_data(uint64_t seed, int n, double gamma, double *__restrict__ u_out,
                float *__restrict__ x_out, float *__restrict__ y_out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s = seed*0x2545F4914F6CDD1Dull + (uint64_t)i*0x9E3779B97F4A7C15ull;
  smix(s);
  double v[8], u[8], d[3];
  const double rho = pow(10.0, -2.0 + 4.0*unif(s));
  const double th  = pow(10.0, -3.0 + 5.0*unif(s));
  const double lor = pow(10.0, log10(20.0)*unif(s));
  const double vm  = sqrt(1.0 - 1.0/(lor*lor));
  unit_vector(s, d);
  v[0] = rho; v[1] = vm*d[0]; v[2] = vm*d[1]; v[3] = vm*d[2];
  const double h = 1.0 + gamma/(gamma - 1.0)*th;
  const double sig = pow(10.0, -3.0 + 5.0*unif(s));
  const double bm = sqrt(sig*rho*h);
  unit_vector(s, d);
  v[4] = bm*d[0]; v[5] = bm*d[1]; v[6] = bm*d[2];
  v[7] = rho*th;
  rmhd_prim_to_cons<double>(v, u, gamma);
  for (int q = 0; q < 8; q++) u_out[(size_t)i*8 + q] = u[q];
  float x[NET_F];
  net_features(u, rmhd_guess_a27(u), x);
  for (int q = 0; q < NET_F; q++) x_out[(size_t)i*NET_F + q] = x[q];
  y_out[i] = net_target(u, v, gamma);
}

// features and targets of given (conserved, primitive) pairs -- the states
// of a run, for instance
__global__ __launch_bounds__(256)
void k_net_pairs(const double *__restrict__ uin, const double *__restrict__ vin, int n, double gamma,
                 float *__restrict__ x_out, float *__restrict__ y_out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  double u[8], v[8];
  for (int q = 0; q < 8; q++) { u[q] = uin[(size_t)i*8 + q]; v[q] = vin[(size_t)i*8 + q]; }
  float x[NET_F];
  net_features(u, rmhd_guess_a27(u), x);
  for (int q = 0; q < NET_F; q++) x_out[(size_t)i*NET_F + q] = x[q];
  y_out[i] = net_target(u, v, gamma);
}

constexpr int TB = 64;                 // samples per chunk (one per thread)
constexpr int NTRAIN = NET_TRAIN_END - NET_TRAIN_BEG;

// Forward + backward of the mean-squared error over the samples
// [off, off + nb) (cyclic in n), in FP32.  Workgroup w takes chunks of TB
// samples w, w + gridDim.x, ... and writes its summed gradient (trainable
// range) and loss to partial[w][*]; k_net_reduce adds the workgroups in a
// fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(TB)
void k_net_train(const float *__restrict__ net, const float *__restrict__ x,
                 const float *__restrict__ y, int off, int n, int nb, float *__restrict__ partial)
{
  __shared__ float sA1[NET_H][TB + 1], sA2[NET_H][TB + 1], sD1[NET_H][TB + 1], sD2[NET_H][TB + 1];
  __shared__ float sX[NET_F][TB + 1], sG[TB], sL[TB];
  __shared__ float sGrad[NTRAIN];
  const int t = threadIdx.x;
  for (int p = t; p < NTRAIN; p += TB) sGrad[p] = 0.f;
  float lsum = 0.f;
  const int nchunk = (nb + TB - 1)/TB;
  for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const int b = ch*TB + t;
    const bool act = b < nb;
    const int i = (off + (act ? b : 0)) % n;
    float xn[NET_F], a1[NET_H];
#pragma unroll
    for (int f = 0; f < NET_F; f++) {
      xn[f] = (x[(size_t)i*NET_F + f] - net[NET_MU + f])/net[NET_SD + f];
      sX[f][t] = xn[f];
    }
    const float yt = y[i]/net[NET_YSC];
#pragma unroll
    for (int h = 0; h < NET_H; h++) {
      float z = net[NET_B1 + h];
#pragma unroll
      for (int f = 0; f < NET_F; f++) z += net[NET_W1 + h*NET_F + f]*xn[f];
      a1[h] = fmaxf(z, 0.f);
      sA1[h][t] = a1[h];
    }
    float out = net[NET_B3];
    for (int o = 0; o < NET_H; o++) {
      float z = net[NET_B2 + o];
#pragma unroll
      for (int h = 0; h < NET_H; h++) z += net[NET_W2 + o*NET_H + h]*a1[h];
      const float a2 = fmaxf(z, 0.f);
      sA2[o][t] = a2;
      out += net[NET_W3 + o]*a2;
    }
    const float err = act ? out - yt : 0.f;
    const float g = 2.f*err/(float)nb;
    float d1[NET_H];
#pragma unroll
    for (int h = 0; h < NET_H; h++) d1[h] = 0.f;
    for (int o = 0; o < NET_H; o++) {
      const float d2 = sA2[o][t] > 0.f ? g*net[NET_W3 + o] : 0.f;
      sD2[o][t] = d2;
#pragma unroll
      for (int h = 0; h < NET_H; h++) d1[h] += net[NET_W2 + o*NET_H + h]*d2;
    }
#pragma unroll
    for (int h = 0; h < NET_H; h++) sD1[h][t] = a1[h] > 0.f ? d1[h] : 0.f;
    sG[t] = g;
    sL[t] = err*err/(float)nb;
    __syncthreads();
    for (int p = t; p < NTRAIN; p += TB) {
      const int q = NET_TRAIN_BEG + p;
      float s = 0.f;
      if (q < NET_B1)      { const int h = (q - NET_W1)/NET_F, f = (q - NET_W1)%NET_F;
                             for (int k = 0; k < TB; k++) s += sD1[h][k]*sX[f][k]; }
      else if (q < NET_W2) { const int h = q - NET_B1; for (int k = 0; k < TB; k++) s += sD1[h][k]; }
      else if (q < NET_B2) { const int o = (q - NET_W2)/NET_H, h = (q - NET_W2)%NET_H;
                             for (int k = 0; k < TB; k++) s += sD2[o][k]*sA1[h][k]; }
      else if (q < NET_W3) { const int o = q - NET_B2; for (int k = 0; k < TB; k++) s += sD2[o][k]; }
      else if (q < NET_B3) { const int h = q - NET_W3; for (int k = 0; k < TB; k++) s += sG[k]*sA2[h][k]; }
      else                 { for (int k = 0; k < TB; k++) s += sG[k]; }
      sGrad[p] += s;
    }
    if (t == 0) for (int k = 0; k < TB; k++) lsum += sL[k];
    __syncthreads();
  }
  float *out = partial + (size_t)blockIdx.x*NET_SIZE;
  for (int p = t; p < NTRAIN; p += TB) out[NET_TRAIN_BEG + p] = sGrad[p];
  if (t == 0) out[NET_SIZE - 1] = lsum;
}

// grad[q] = sum over workgroups of partial[w][q] (trainable range and the
// loss slot, the last entry)
__global__ void k_net_reduce(const float *__restrict__ partial, int nwg, float *__restrict__ grad)
{
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  int q;
  if (p < NTRAIN) q = NET_TRAIN_BEG + p;
  else if (p == NTRAIN) q = NET_SIZE - 1;
  else return;
  float s = 0.f;
  for (int w = 0; w < nwg; w++) s += partial[(size_t)w*NET_SIZE + q];
  grad[q] = s;
}

__global__ void k_net_zero(float *__restrict__ a, int n)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.f;
}

// Adam (Kingma & Ba 2015) on the trainable range
__global__ void k_net_adam(float *__restrict__ net, const float *__restrict__ grad,
                           float *__restrict__ m, float *__restrict__ v, float lr, float c1, float c2)
{
  const int q = NET_TRAIN_BEG + blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= NET_TRAIN_END) return;
  const float g = grad[q];
  m[q] = 0.9f*m[q] + 0.1f*g;
  v[q] = 0.999f*v[q] + 0.001f*g*g;
  net[q] -= lr*(m[q]/c1)/(sqrtf(v[q]/c2) + 1e-8f);
}

__device__ __forceinline__ uint16_t bf16_bits(float f)
{
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

// the matrix-core form of the weight vector (layout in c2p_net.hpp)
__global__ __launch_bounds__(256)
void k_net_pack(const float *__restrict__ w, NetPack *__restrict__ P)
{
  const int t = threadIdx.x;
  for (int e = t; e < 2*2*64*8; e += blockDim.x) {
    const int j = e & 7, lane = (e >> 3) & 63, mt = (e >> 9) & 1, pass = e >> 10;
    const int row = 32*mt + (lane & 31), hi = lane >> 5;
    const float val = hi == pass ? w[NET_W1 + row*NET_F + j] : (j == 0 ? w[NET_B1 + row] : 0.f);
    P->l1[pass][mt][lane][j] = bf16_bits(val);
  }
  for (int e = t; e < 2*5*64*8; e += blockDim.x) {
    const int j = e & 7, lane = (e >> 3) & 63, s = (e >> 9)%5, mt = (e >> 9)/5;
    const int row = 32*mt + (lane & 31), hi = lane >> 5;
    const float val = s < 4 ? w[NET_W2 + row*NET_H + net_unit(s, hi, j)]
                            : (hi == 0 && j == 0 ? w[NET_B2 + row] : 0.f);
    P->l2[mt][s][lane][j] = bf16_bits(val);
  }
  for (int e = t; e < 2*5*64*8; e += blockDim.x) {
    const int j = e & 7, lane = (e >> 3) & 63, s = (e >> 9)%5, pass = (e >> 9)/5;
    const int row = lane & 31, hi = lane >> 5;
    float val = 0.f;
    if (row == (pass == 0 ? 0 : 4))
      val = s < 4 ? w[NET_W3 + net_unit(s, hi, j)] : (hi == 0 && j == 0 ? w[NET_B3] : 0.f);
    P->l3[pass][s][lane][j] = bf16_bits(val);
  }
  if (t < NET_F) {
    P->mu[t] = w[NET_MU + t];
    P->rsd[t] = 1.f/w[NET_SD + t];
  }
  if (t == 0) P->ysc = w[NET_YSC];
}

// matrix-core inference: the correction ysc * net(x) of every sample
__global__ __launch_bounds__(256)
void k_net_infer(const NetPack *__restrict__ P, const float *__restrict__ x, int n,
                 float *__restrict__ out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  float xs[NET_F];
#pragma unroll
  for (int q = 0; q < NET_F; q++) xs[q] = i < n ? x[(size_t)i*NET_F + q] : 0.f;
  const float o = net_eval64(P, xs);      // every lane of the wavefront
  if (i < n) out[i] = o;
}

// batched energy solve of RMHD con2prim from the A27 guess (P == nullptr)
// or from the learned guess with the A27 fallback
__global__ __launch_bounds__(256)
void k_net_solve(const NetPack *__restrict__ P, const double *__restrict__ uin, int n, double gamma,
                 double *__restrict__ vout, int *__restrict__ passes, uint32_t *__restrict__ errs,
                 int *__restrict__ restarts)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  const bool live = i < n;
  double u[8], v[8];
  for (int q = 0; q < 8; q++) u[q] = uin[(size_t)(live ? i : 0)*8 + q];
  for (int q = 0; q < 8; q++) v[q] = u[q];
  int it = 0, again = 0;
  uint32_t e;
  if (P) {
    const double wa = rmhd_guess_a27(u);
    float x[NET_F];
    net_features(u, wa, x);
    const float corr = net_eval64(P, x);          // every lane of the wavefront
    e = rmhd_energy_solve_guided(u, v, gamma, wa, corr, &it, &again);
  } else {
    e = rmhd_energy_solve(u, v, gamma, &it);
  }
  if (!live) return;
  for (int q = 0; q < 8; q++) vout[(size_t)i*8 + q] = v[q];
  passes[i] = it;
  errs[i] = e;
  if (restarts) restarts[i] = again;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int alloc(size_t bytes) { return hipMalloc(&p, bytes) == hipSuccess ? 0 : GPL_ENOMEM; }
  template <typename T> T *as() { return (T *)p; }
};

int set_device(int device)
{
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return GPL_EHIP;
  return hipSetDevice(device) == hipSuccess ? 0 : GPL_EHIP;
}

// weight vector (host) -> NetPack (device), through a device copy of the vector
int pack_weights(const float *weights, float *dev_w, NetPack *dev_p, hipStream_t s)
{
  if (hipMemcpyAsync(dev_w, weights, NET_SIZE*sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess)
    return GPL_EHIP;
  hipLaunchKernelGGL(k_net_pack, dim3(1), dim3(256), 0, s, dev_w, dev_p);
  return hipGetLastError() == hipSuccess ? 0 : GPL_EHIP;
}

constexpr int NWG = 64;   // training workgroups (partial gradients)

// Train on (x, y) resident on the device: normalisation from the data (on
// the host), Glorot-scaled initial weights from a seeded stream, Adam with
// the rate cut by 3 for the last quarter of the steps.
int fit_on_device(uint64_t seed, const float *dx, const float *dy, int n, int steps, int batch,
                  float lr, float *weights_out, float *loss_out)
{
  std::vector<float> hx((size_t)n*NET_F), hy(n);
  if (hipMemcpy(hx.data(), dx, hx.size()*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hy.data(), dy, hy.size()*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  std::vector<float> net(NET_SIZE, 0.f);
  for (int f = 0; f < NET_F; f++) {
    double s = 0, s2 = 0;
    for (int i = 0; i < n; i++) {
      const double a = hx[(size_t)i*NET_F + f];
      s += a; s2 += a*a;
    }
    const double mu = s/n, var = s2/n - mu*mu;
    net[NET_MU + f] = (float)mu;
    net[NET_SD + f] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
  }
  {
    double s2 = 0;
    for (int i = 0; i < n; i++) s2 += (double)hy[i]*hy[i];
    net[NET_YSC] = (float)(s2 > 0 ? std::sqrt(s2/n) : 1.0);
  }
  uint64_t st = seed ^ 0x5DEECE66DULL;
  auto nrm = [&st]() {
    const double a = (double)(smix(st) >> 11)*0x1.0p-53, b = (double)(smix(st) >> 11)*0x1.0p-53;
    return std::sqrt(-2.0*std::log(a + 1e-300))*std::cos(6.283185307179586*b);
  };
  // He-scaled for ReLU layers
  for (int q = 0; q < NET_H*NET_F; q++) net[NET_W1 + q] = (float)(nrm()*std::sqrt(2.0/NET_F));
  for (int q = 0; q < NET_H*NET_H; q++) net[NET_W2 + q] = (float)(nrm()*std::sqrt(2.0/NET_H));
  for (int q = 0; q < NET_H; q++) net[NET_W3 + q] = (float)(nrm()/std::sqrt((double)NET_H));
  for (int q = 0; q < NET_H; q++) net[NET_B1 + q] = net[NET_B2 + q] = 0.01f;
  DevBuf w, g, m, v, part;
  if (w.alloc(NET_SIZE*sizeof(float)) || g.alloc(NET_SIZE*sizeof(float)) ||
      m.alloc(NET_SIZE*sizeof(float)) || v.alloc(NET_SIZE*sizeof(float)) ||
      part.alloc((size_t)NWG*NET_SIZE*sizeof(float)))
    return GPL_ENOMEM;
  if (hipMemcpy(w.p, net.data(), NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  const int zb = (NET_SIZE + 255)/256;
  hipLaunchKernelGGL(k_net_zero, dim3(zb), dim3(256), 0, 0, m.as<float>(), NET_SIZE);
  hipLaunchKernelGGL(k_net_zero, dim3(zb), dim3(256), 0, 0, v.as<float>(), NET_SIZE);
  const int nwg = std::min(NWG, (batch + TB - 1)/TB);
  float c1 = 1.f, c2 = 1.f, last = 0.f;
  for (int s = 0; s < steps; s++) {
    const float rate = s < (3*steps)/4 ? lr : lr/3.f;
    c1 *= 0.9f; c2 *= 0.999f;
    hipLaunchKernelGGL(k_net_train, dim3(nwg), dim3(TB), 0, 0, w.as<float>(), dx, dy,
                       (int)(((long long)s*batch) % n), n, batch, part.as<float>());
    hipLaunchKernelGGL(k_net_reduce, dim3((NTRAIN + 1 + 255)/256), dim3(256), 0, 0,
                       part.as<float>(), nwg, g.as<float>());
    hipLaunchKernelGGL(k_net_adam, dim3((NTRAIN + 255)/256), dim3(256), 0, 0, w.as<float>(),
                       g.as<float>(), m.as<float>(), v.as<float>(), rate, 1.f - c1, 1.f - c2);
    if (s == steps - 1 &&
        hipMemcpy(&last, g.as<float>() + NET_SIZE - 1, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return GPL_EHIP;
  }
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(weights_out, w.p, NET_SIZE*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  weights_out[NET_SIZE - 1] = 0.f;
  if (loss_out) *loss_out = last;
  return 0;
}

}  // namespace

// the engine's upload of a weight vector into its NetPack
int net_pack_upload(const float *weights, float *dev_w, NetPack *dev_p, hipStream_t s)
{
  return pack_weights(weights, dev_w, dev_p, s);
}

}  // namespace gpl

using namespace gpl;

extern "C" int gpl_c2p_net_size(void) { return NET_SIZE; }

extern "C" int gpl_c2p_net_make_data(int device, uint64_t seed, int n, double gamma, double *u_out,
                                     float *x_out, float *y_out)
{
  if (n < 1 || !u_out || !x_out || !y_out) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf u, x, y;
  if (u.alloc((size_t)n*8*sizeof(double)) || x.alloc((size_t)n*NET_F*sizeof(float)) ||
      y.alloc((size_t)n*sizeof(float)))
    return GPL_ENOMEM;
  hipLaunchKernelGGL(k_net_data, dim3((n + 255)/256), dim3(256), 0, 0, seed, n, gamma,
                     u.as<double>(), x.as<float>(), y.as<float>());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(u_out, u.p, (size_t)n*8*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(x_out, x.p, (size_t)n*NET_F*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(y_out, y.p, (size_t)n*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

extern "C" int gpl_c2p_net_pairs(int device, const double *u, const double *v, int n, double gamma,
                                 float *x_out, float *y_out)
{
  if (n < 1 || !u || !v || !x_out || !y_out) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf du, dv, x, y;
  if (du.alloc((size_t)n*8*sizeof(double)) || dv.alloc((size_t)n*8*sizeof(double)) ||
      x.alloc((size_t)n*NET_F*sizeof(float)) || y.alloc((size_t)n*sizeof(float)))
    return GPL_ENOMEM;
  if (hipMemcpy(du.p, u, (size_t)n*8*sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dv.p, v, (size_t)n*8*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipLaunchKernelGGL(k_net_pairs, dim3((n + 255)/256), dim3(256), 0, 0, du.as<double>(),
                     dv.as<double>(), n, gamma, x.as<float>(), y.as<float>());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(x_out, x.p, (size_t)n*NET_F*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(y_out, y.p, (size_t)n*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

extern "C" int gpl_c2p_net_grad(int device, const float *weights, const float *x, const float *y,
                                int n, float *grad_out, float *loss_out)
{
  if (!weights || !x || !y || !grad_out || n < 1) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf w, dx, dy, g, part;
  if (w.alloc(NET_SIZE*sizeof(float)) || dx.alloc((size_t)n*NET_F*sizeof(float)) ||
      dy.alloc((size_t)n*sizeof(float)) || g.alloc(NET_SIZE*sizeof(float)) ||
      part.alloc((size_t)NWG*NET_SIZE*sizeof(float)))
    return GPL_ENOMEM;
  if (hipMemcpy(w.p, weights, NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx.p, x, (size_t)n*NET_F*sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy.p, y, (size_t)n*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  const int nwg = std::min(NWG, (n + TB - 1)/TB);
  hipLaunchKernelGGL(k_net_zero, dim3((NET_SIZE + 255)/256), dim3(256), 0, 0, g.as<float>(), NET_SIZE);
  hipLaunchKernelGGL(k_net_train, dim3(nwg), dim3(TB), 0, 0, w.as<float>(), dx.as<float>(),
                     dy.as<float>(), 0, n, n, part.as<float>());
  hipLaunchKernelGGL(k_net_reduce, dim3((NTRAIN + 1 + 255)/256), dim3(256), 0, 0, part.as<float>(),
                     nwg, g.as<float>());
  float loss = 0.f;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(grad_out, g.p, NET_SIZE*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  loss = grad_out[NET_SIZE - 1];
  grad_out[NET_SIZE - 1] = 0.f;
  if (loss_out) *loss_out = loss;
  return 0;
}

extern "C" int gpl_c2p_net_infer(int device, const float *weights, const float *x, int n, float *y_out)
{
  if (!weights || !x || !y_out || n < 1) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf w, pk, dx, dy;
  if (w.alloc(NET_SIZE*sizeof(float)) || pk.alloc(sizeof(NetPack)) ||
      dx.alloc((size_t)n*NET_F*sizeof(float)) || dy.alloc((size_t)n*sizeof(float)))
    return GPL_ENOMEM;
  if (pack_weights(weights, w.as<float>(), pk.as<NetPack>(), 0) ||
      hipMemcpy(dx.p, x, (size_t)n*NET_F*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipLaunchKernelGGL(k_net_infer, dim3((n + 255)/256), dim3(256), 0, 0, pk.as<NetPack>(),
                     dx.as<float>(), n, dy.as<float>());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(y_out, dy.p, (size_t)n*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

extern "C" int gpl_c2p_net_fit(int device, uint64_t seed, const float *x, const float *y, int n,
                               int steps, int batch, float lr, float *weights_out, float *loss_out)
{
  if (!x || !y || !weights_out || n < TB || steps < 1 || batch < 1 || batch > n) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  // a seeded shuffle, so that cyclic mini-batches of a run's states are not
  // neighbouring zones
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  uint64_t st = seed*0x2545F4914F6CDD1Dull + 1;
  for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int)(smix(st) % (uint64_t)(i + 1))]);
  std::vector<float> hx((size_t)n*NET_F), hy(n);
  for (int i = 0; i < n; i++) {
    std::memcpy(&hx[(size_t)i*NET_F], x + (size_t)perm[i]*NET_F, NET_F*sizeof(float));
    hy[i] = y[perm[i]];
  }
  DevBuf dx, dy;
  if (dx.alloc(hx.size()*sizeof(float)) || dy.alloc(hy.size()*sizeof(float))) return GPL_ENOMEM;
  if (hipMemcpy(dx.p, hx.data(), hx.size()*sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy.p, hy.data(), hy.size()*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  return fit_on_device(seed, dx.as<float>(), dy.as<float>(), n, steps, batch, lr, weights_out, loss_out);
}

extern "C" int gpl_c2p_net_train(int device, uint64_t seed, int nsamples, int steps, int batch,
                                 float lr, double gamma, float *weights_out, float *loss_out)
{
  if (!weights_out || nsamples < TB || steps < 1 || batch < 1 || batch > nsamples) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  const int n = nsamples;
  DevBuf u, x, y;
  if (u.alloc((size_t)n*8*sizeof(double)) || x.alloc((size_t)n*NET_F*sizeof(float)) ||
      y.alloc((size_t)n*sizeof(float)))
    return GPL_ENOMEM;
  hipLaunchKernelGGL(k_net_data, dim3((n + 255)/256), dim3(256), 0, 0, seed, n, gamma,
                     u.as<double>(), x.as<float>(), y.as<float>());
  if (hipDeviceSynchronize() != hipSuccess) return GPL_EHIP;
  return fit_on_device(seed, x.as<float>(), y.as<float>(), n, steps, batch, lr, weights_out, loss_out);
}

extern "C" int gpl_c2p_solve(int device, const float *weights, const double *u, int n, double gamma,
                             double *v_out, int *passes_out, uint32_t *err_out, double *ms_out,
                             int *restart_out)
{
  if (!u || !v_out || !passes_out || !err_out || n < 1) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf w, pk, du, dv, dp, de, dr;
  if ((weights && (w.alloc(NET_SIZE*sizeof(float)) || pk.alloc(sizeof(NetPack)))) ||
      du.alloc((size_t)n*8*sizeof(double)) || dv.alloc((size_t)n*8*sizeof(double)) ||
      dp.alloc((size_t)n*sizeof(int)) || de.alloc((size_t)n*sizeof(uint32_t)) ||
      (restart_out && dr.alloc((size_t)n*sizeof(int))))
    return GPL_ENOMEM;
  if ((weights && pack_weights(weights, w.as<float>(), pk.as<NetPack>(), 0)) ||
      hipMemcpy(du.p, u, (size_t)n*8*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return GPL_EHIP;
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return GPL_EHIP; }
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k_net_solve, dim3((n + 255)/256), dim3(256), 0, 0,
                     weights ? pk.as<NetPack>() : nullptr, du.as<double>(), n, gamma, dv.as<double>(),
                     dp.as<int>(), de.as<uint32_t>(), restart_out ? dr.as<int>() : nullptr);
  (void)hipEventRecord(e1, 0);
  const bool ok = hipEventSynchronize(e1) == hipSuccess;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!ok || hipMemcpy(v_out, dv.p, (size_t)n*8*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(passes_out, dp.p, (size_t)n*sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(err_out, de.p, (size_t)n*sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  if (restart_out) {
    if (!weights) {
      for (int i = 0; i < n; i++) restart_out[i] = 0;
    } else if (hipMemcpy(restart_out, dr.p, (size_t)n*sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      return GPL_EHIP;
    }
  }
  if (ms_out) *ms_out = ms;
  return 0;
}

// ---------------------------------------------------------------------------
// gpl_c2p_solve without allocation: the states, results and weights are
// staged in one file-scope device array instead of hipMalloc'd buffers.  The
// host reaches the array only through the address hipGetSymbolAddress
// returns for the current device (each device holds its own copy).
// ---------------------------------------------------------------------------
namespace gpl {
namespace {

constexpr size_t kNoAllocMax = GPL_C2P_NOALLOC_MAX;
constexpr size_t kScrU = 0;                                        // [max][8] double
constexpr size_t kScrV = kScrU + kNoAllocMax*8*sizeof(double);     // [max][8] double
constexpr size_t kScrP = kScrV + kNoAllocMax*8*sizeof(double);     // [max] int
constexpr size_t kScrE = kScrP + kNoAllocMax*sizeof(int);          // [max] uint32
constexpr size_t kScrR = kScrE + kNoAllocMax*sizeof(int);          // [max] int
constexpr size_t kScrW = kScrR + kNoAllocMax*sizeof(int);          // [NET_SIZE] float
constexpr size_t kScrK = kScrW + (NET_SIZE*sizeof(float) + 255)/256*256;   // NetPack
constexpr size_t kScrBytes = kScrK + (sizeof(NetPack) + 255)/256*256;
static_assert(kScrV % 256 == 0 && kScrP % 256 == 0 && kScrE % 256 == 0 && kScrR % 256 == 0 &&
              kScrW % 256 == 0 && kScrK % 256 == 0, "scratch regions are 256-byte aligned");

__device__ __align__(256) unsigned char scratch[kScrBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
unsigned char *scratch_dev[kMaxDevices];   // the array's address on each device

// the scratch address on the current device (looked up once per device)
int scratch_on(int device, unsigned char **out)
{
  if (device < 0 || device >= kMaxDevices) return GPL_EINVAL;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GPL_EHIP;
    scratch_dev[device] = (unsigned char *)p;
  }
  *out = scratch_dev[device];
  return 0;
}

}  // namespace
}  // namespace gpl

extern "C" int gpl_c2p_solve_noalloc(int device, const float *weights, const double *u, int n,
                                     double gamma, double *v_out, int *passes_out, uint32_t *err_out,
                                     int *restart_out)
{
  using namespace gpl;
  if (!u || !v_out || !passes_out || !err_out || n < 1) return GPL_EINVAL;
  if ((size_t)n > kNoAllocMax) return GPL_ESCRATCH;
  if (int rc = set_device(device)) return rc;
  unsigned char *base = nullptr;
  if (int rc = scratch_on(device, &base)) return rc;
  double *du = (double *)(base + kScrU), *dv = (double *)(base + kScrV);
  int *dp = (int *)(base + kScrP), *dr = (int *)(base + kScrR);
  uint32_t *de = (uint32_t *)(base + kScrE);
  float *dw = (float *)(base + kScrW);
  NetPack *dk = (NetPack *)(base + kScrK);
  if ((weights && pack_weights(weights, dw, dk, 0)) ||
      hipMemcpy(du, u, (size_t)n*8*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  // every output slot below n is written by the kernel: nothing to clear
  hipLaunchKernelGGL(k_net_solve, dim3((n + 255)/256), dim3(256), 0, 0, weights ? dk : nullptr, du, n,
                     gamma, dv, dp, de, (restart_out && weights) ? dr : nullptr);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(v_out, dv, (size_t)n*8*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(passes_out, dp, (size_t)n*sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(err_out, de, (size_t)n*sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  if (restart_out) {
    if (!weights) {
      for (int i = 0; i < n; i++) restart_out[i] = 0;
    } else if (hipMemcpy(restart_out, dr, (size_t)n*sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      return GPL_EHIP;
    }
  }
  return 0;
}
