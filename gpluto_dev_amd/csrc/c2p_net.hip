// Learned con2prim initial guess for RMHD: training-data generation, FP32
// forward/backward training kernels with Adam, matrix-core inference, and a
// batched Newton solve that reports its passes -- behind the C ABI.
#include "../../include/gpluto_amd.h"
#include "c2p_net.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace gpl {

namespace {

// splitmix64 (Steele, Lea & Flood 2014): a counter-based stream per sample
__device__ __forceinline__ uint64_t smix(uint64_t &s)
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

// (conserved, primitive) pairs made by the method itself: a primitive state
// drawn over rho in [1e-2,1e2], p/rho in [1e-3,1e2], Lorentz factor in
// [1,20], sigma = B^2/(rho h) in [1e-3,1e2] (log-uniform, isotropic
// directions), mapped by PrimToCons; the target is log(W/W_A27).
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
  const double g2 = 1.0/(1.0 - (v[1]*v[1] + v[2]*v[2] + v[3]*v[3]));
  const double W = rho*h*g2;
  for (int q = 0; q < 8; q++) u_out[(size_t)i*8 + q] = u[q];
  float x[4];
  net_features(u, x);
  for (int q = 0; q < 4; q++) x_out[(size_t)i*4 + q] = x[q];
  y_out[i] = (float)log(W/rmhd_guess_a27(u));
}

constexpr int TB = 128;   // samples per training workgroup

// Forward + backward of the mean-squared error over a mini-batch of nb
// samples starting at 'off' (cyclic in n); gradients (of the trainable range)
// accumulate into grad, the batch loss into loss[0].
__global__ __launch_bounds__(TB)
void k_net_train(const float *__restrict__ net, const float *__restrict__ x,
                 const float *__restrict__ y, int off, int n, int nb, float *__restrict__ grad,
                 float *__restrict__ loss)
{
  __shared__ float sA1[TB][NET_H + 1], sA2[TB][NET_H + 1], sD1[TB][NET_H + 1], sD2[TB][NET_H + 1];
  __shared__ float sX[TB][4], sG[TB], sL[TB];
  const int t = threadIdx.x, b = blockIdx.x*TB + t;
  const bool act = b < nb;
  const int i = (off + (act ? b : 0)) % n;
  float xn[4], a1[NET_H], a2[NET_H];
  for (int f = 0; f < 4; f++) xn[f] = (x[(size_t)i*4 + f] - net[NET_MU + f])/net[NET_SD + f];
  const float yt = y[i]/net[NET_YSC];
  for (int h = 0; h < NET_H; h++) {
    float z = net[NET_B1 + h];
    for (int f = 0; f < 4; f++) z += net[NET_W1 + h*4 + f]*xn[f];
    a1[h] = tanhf(z);
  }
  float out = net[NET_B3];
  for (int o = 0; o < NET_H; o++) {
    float z = net[NET_B2 + o];
    for (int h = 0; h < NET_H; h++) z += net[NET_W2 + o*NET_H + h]*a1[h];
    a2[o] = tanhf(z);
    out += net[NET_W3 + o]*a2[o];
  }
  const float err = act ? out - yt : 0.f;
  const float g = 2.f*err/(float)nb;
  float d2[NET_H];
  for (int o = 0; o < NET_H; o++) {
    d2[o] = g*net[NET_W3 + o]*(1.f - a2[o]*a2[o]);
    sD2[t][o] = d2[o];
    sA2[t][o] = a2[o];
    sA1[t][o] = a1[o];
  }
  for (int h = 0; h < NET_H; h++) {
    float s = 0.f;
    for (int o = 0; o < NET_H; o++) s += net[NET_W2 + o*NET_H + h]*d2[o];
    sD1[t][h] = s*(1.f - a1[h]*a1[h]);
  }
  for (int f = 0; f < 4; f++) sX[t][f] = xn[f];
  sG[t] = g;
  sL[t] = err*err/(float)nb;
  __syncthreads();
  for (int p = t; p < NET_TRAIN_END - NET_TRAIN_BEG; p += TB) {
    const int q = NET_TRAIN_BEG + p;
    float s = 0.f;
    if (q < NET_B1)      { const int h = (q - NET_W1)/4, f = (q - NET_W1)%4;
                           for (int k = 0; k < TB; k++) s += sD1[k][h]*sX[k][f]; }
    else if (q < NET_W2) { const int h = q - NET_B1; for (int k = 0; k < TB; k++) s += sD1[k][h]; }
    else if (q < NET_B2) { const int o = (q - NET_W2)/NET_H, h = (q - NET_W2)%NET_H;
                           for (int k = 0; k < TB; k++) s += sD2[k][o]*sA1[k][h]; }
    else if (q < NET_W3) { const int o = q - NET_B2; for (int k = 0; k < TB; k++) s += sD2[k][o]; }
    else if (q < NET_B3) { const int h = q - NET_W3; for (int k = 0; k < TB; k++) s += sG[k]*sA2[k][h]; }
    else                 { for (int k = 0; k < TB; k++) s += sG[k]; }
    atomicAdd(&grad[q], s);
  }
  if (t == 0) {
    float s = 0.f;
    for (int k = 0; k < TB; k++) s += sL[k];
    atomicAdd(loss, s);
  }
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

// matrix-core inference: one wavefront per 64 samples
__global__ __launch_bounds__(256)
void k_net_infer(const float *__restrict__ net, const float *__restrict__ x, int n,
                 float *__restrict__ out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  float xs[4] = {0.f, 0.f, 0.f, 0.f};
  if (i < n) for (int q = 0; q < 4; q++) xs[q] = x[(size_t)i*4 + q];
  const float o = net_eval64(net, xs);
  if (i < n) out[i] = o;
}

// batched energy solve of RMHD con2prim from the A27 guess (net == nullptr)
// or from the learned guess with the A27 fallback
__global__ __launch_bounds__(256)
void k_net_solve(const float *__restrict__ net, const double *__restrict__ uin, int n, double gamma,
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
  if (net) {
    float x[4];
    net_features(u, x);
    const float pred = net_eval64(net, x);          // every lane of the wavefront
    e = rmhd_energy_solve_guided(u, v, gamma, pred, net[NET_YSC], &it, &again);
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

}  // namespace
}  // namespace gpl

using namespace gpl;

extern "C" int gpl_c2p_net_size(void) { return NET_SIZE; }

extern "C" int gpl_c2p_net_make_data(int device, uint64_t seed, int n, double gamma, double *u_out,
                                     float *x_out, float *y_out)
{
  if (n < 1 || !u_out || !x_out || !y_out) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf u, x, y;
  if (u.alloc((size_t)n*8*sizeof(double)) || x.alloc((size_t)n*4*sizeof(float)) ||
      y.alloc((size_t)n*sizeof(float)))
    return GPL_ENOMEM;
  hipLaunchKernelGGL(k_net_data, dim3((n + 255)/256), dim3(256), 0, 0, seed, n, gamma,
                     u.as<double>(), x.as<float>(), y.as<float>());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(u_out, u.p, (size_t)n*8*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(x_out, x.p, (size_t)n*4*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(y_out, y.p, (size_t)n*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

extern "C" int gpl_c2p_net_grad(int device, const float *weights, const float *x, const float *y,
                                int n, float *grad_out, float *loss_out)
{
  if (!weights || !x || !y || !grad_out || n < 1) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf w, dx, dy, g, l;
  if (w.alloc(NET_SIZE*sizeof(float)) || dx.alloc((size_t)n*4*sizeof(float)) ||
      dy.alloc((size_t)n*sizeof(float)) || g.alloc(NET_SIZE*sizeof(float)) || l.alloc(sizeof(float)))
    return GPL_ENOMEM;
  if (hipMemcpy(w.p, weights, NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx.p, x, (size_t)n*4*sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy.p, y, (size_t)n*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipLaunchKernelGGL(k_net_zero, dim3((NET_SIZE + 255)/256), dim3(256), 0, 0, g.as<float>(), NET_SIZE);
  hipLaunchKernelGGL(k_net_zero, dim3(1), dim3(64), 0, 0, l.as<float>(), 1);
  hipLaunchKernelGGL(k_net_train, dim3((n + TB - 1)/TB), dim3(TB), 0, 0, w.as<float>(),
                     dx.as<float>(), dy.as<float>(), 0, n, n, g.as<float>(), l.as<float>());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(grad_out, g.p, NET_SIZE*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  if (loss_out && hipMemcpy(loss_out, l.p, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

extern "C" int gpl_c2p_net_infer(int device, const float *weights, const float *x, int n, float *y_out)
{
  if (!weights || !x || !y_out || n < 1) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf w, dx, dy;
  if (w.alloc(NET_SIZE*sizeof(float)) || dx.alloc((size_t)n*4*sizeof(float)) ||
      dy.alloc((size_t)n*sizeof(float)))
    return GPL_ENOMEM;
  if (hipMemcpy(w.p, weights, NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx.p, x, (size_t)n*4*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipLaunchKernelGGL(k_net_infer, dim3((n + 255)/256), dim3(256), 0, 0, w.as<float>(),
                     dx.as<float>(), n, dy.as<float>());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(y_out, dy.p, (size_t)n*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

extern "C" int gpl_c2p_net_train(int device, uint64_t seed, int nsamples, int steps, int batch,
                                 float lr, double gamma, float *weights_out, float *loss_out)
{
  if (!weights_out || nsamples < TB || steps < 1 || batch < 1 || batch > nsamples) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  const int n = nsamples;
  DevBuf u, x, y, w, g, m, v, l;
  if (u.alloc((size_t)n*8*sizeof(double)) || x.alloc((size_t)n*4*sizeof(float)) ||
      y.alloc((size_t)n*sizeof(float)) || w.alloc(NET_SIZE*sizeof(float)) ||
      g.alloc(NET_SIZE*sizeof(float)) || m.alloc(NET_SIZE*sizeof(float)) ||
      v.alloc(NET_SIZE*sizeof(float)) || l.alloc(sizeof(float)))
    return GPL_ENOMEM;
  hipLaunchKernelGGL(k_net_data, dim3((n + 255)/256), dim3(256), 0, 0, seed, n, gamma,
                     u.as<double>(), x.as<float>(), y.as<float>());
  // normalisation of the inputs and the target from the data
  std::vector<float> hx((size_t)n*4), hy(n);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(hx.data(), x.p, hx.size()*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hy.data(), y.p, hy.size()*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  std::vector<float> net(NET_SIZE, 0.f);
  for (int f = 0; f < 4; f++) {
    double s = 0, s2 = 0;
    for (int i = 0; i < n; i++) { s += hx[(size_t)i*4 + f]; s2 += (double)hx[(size_t)i*4 + f]*hx[(size_t)i*4 + f]; }
    const double mu = s/n, var = s2/n - mu*mu;
    net[NET_MU + f] = (float)mu;
    net[NET_SD + f] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
  }
  {
    double s2 = 0;
    for (int i = 0; i < n; i++) s2 += (double)hy[i]*hy[i];
    net[NET_YSC] = (float)(s2 > 0 ? std::sqrt(s2/n) : 1.0);
  }
  // Glorot-scaled initial weights from a host splitmix64 stream
  uint64_t st = seed ^ 0x5DEECE66DULL;
  auto nrm = [&st]() {
    auto nx = [&st]() {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27))*0x94D049BB133111EBull;
      return (double)((z ^ (z >> 31)) >> 11)*0x1.0p-53;
    };
    const double a = nx(), b = nx();
    return std::sqrt(-2.0*std::log(a + 1e-300))*std::cos(6.283185307179586*b);
  };
  for (int q = 0; q < NET_H*4; q++) net[NET_W1 + q] = (float)(nrm()/2.0);
  for (int q = 0; q < NET_H*NET_H; q++) net[NET_W2 + q] = (float)(nrm()/std::sqrt((double)NET_H));
  for (int q = 0; q < NET_H; q++) net[NET_W3 + q] = (float)(nrm()/std::sqrt((double)NET_H));
  if (hipMemcpy(w.p, net.data(), NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipLaunchKernelGGL(k_net_zero, dim3((NET_SIZE + 255)/256), dim3(256), 0, 0, m.as<float>(), NET_SIZE);
  hipLaunchKernelGGL(k_net_zero, dim3((NET_SIZE + 255)/256), dim3(256), 0, 0, v.as<float>(), NET_SIZE);
  const int nq = NET_TRAIN_END - NET_TRAIN_BEG;
  float c1 = 1.f, c2 = 1.f, last = 0.f;
  for (int s = 0; s < steps; s++) {
    const float rate = s < (3*steps)/4 ? lr : lr/3.f;
    c1 *= 0.9f; c2 *= 0.999f;
    hipLaunchKernelGGL(k_net_zero, dim3((NET_SIZE + 255)/256), dim3(256), 0, 0, g.as<float>(), NET_SIZE);
    hipLaunchKernelGGL(k_net_zero, dim3(1), dim3(64), 0, 0, l.as<float>(), 1);
    hipLaunchKernelGGL(k_net_train, dim3((batch + TB - 1)/TB), dim3(TB), 0, 0, w.as<float>(),
                       x.as<float>(), y.as<float>(), (int)(((long long)s*batch) % n), n, batch,
                       g.as<float>(), l.as<float>());
    hipLaunchKernelGGL(k_net_adam, dim3((nq + 255)/256), dim3(256), 0, 0, w.as<float>(),
                       g.as<float>(), m.as<float>(), v.as<float>(), rate, 1.f - c1, 1.f - c2);
    if (s == steps - 1 &&
        hipMemcpy(&last, l.p, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return GPL_EHIP;
  }
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(weights_out, w.p, NET_SIZE*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  if (loss_out) *loss_out = last;
  return 0;
}

extern "C" int gpl_c2p_solve(int device, const float *weights, const double *u, int n, double gamma,
                             double *v_out, int *passes_out, uint32_t *err_out, double *ms_out,
                             int *restart_out)
{
  if (!u || !v_out || !passes_out || !err_out || n < 1) return GPL_EINVAL;
  if (int rc = set_device(device)) return rc;
  DevBuf w, du, dv, dp, de, dr;
  if ((weights && w.alloc(NET_SIZE*sizeof(float))) || du.alloc((size_t)n*8*sizeof(double)) ||
      dv.alloc((size_t)n*8*sizeof(double)) || dp.alloc((size_t)n*sizeof(int)) ||
      de.alloc((size_t)n*sizeof(uint32_t)) || (restart_out && dr.alloc((size_t)n*sizeof(int))))
    return GPL_ENOMEM;
  if ((weights && hipMemcpy(w.p, weights, NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(du.p, u, (size_t)n*8*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k_net_solve, dim3((n + 255)/256), dim3(256), 0, 0,
                     weights ? w.as<float>() : nullptr, du.as<double>(), n, gamma, dv.as<double>(),
                     dp.as<int>(), de.as<uint32_t>(), restart_out ? dr.as<int>() : nullptr);
  hipEventRecord(e1, 0);
  const bool ok = hipEventSynchronize(e1) == hipSuccess;
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
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
constexpr size_t kScrR = kScrE + kNoAllocMax*sizeof(uint32_t);     // [max] int
constexpr size_t kScrW = kScrR + kNoAllocMax*sizeof(int);          // [NET_SIZE] float
constexpr size_t kScrBytes = kScrW + (NET_SIZE*sizeof(float) + 255)/256*256;
static_assert(kScrV % 256 == 0 && kScrP % 256 == 0 && kScrE % 256 == 0 && kScrR % 256 == 0 &&
              kScrW % 256 == 0, "scratch regions are 256-byte aligned");

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
  if ((weights && hipMemcpy(dw, weights, NET_SIZE*sizeof(float), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(du, u, (size_t)n*8*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  // every output slot below n is written by the kernel: nothing to clear
  hipLaunchKernelGGL(k_net_solve, dim3((n + 255)/256), dim3(256), 0, 0, weights ? dw : nullptr, du, n,
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
