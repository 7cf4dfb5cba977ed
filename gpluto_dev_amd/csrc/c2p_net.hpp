// Learned initial guess for the RMHD con2prim Newton solve.
//
// A 4-32-32-1 tanh network maps scale-free invariants of the conserved state
// to a correction of the Eq. (A27) guess of RMHD_EnergySolve
// (Src/RMHD/rmhd_energy_solve.cpp:86-96): W_guess = W_A27 * exp(ysc * net(x)).
// Inference runs on the matrix cores: one wavefront evaluates 32 states per
// pass with v_mfma_f32_32x32x8_f16 (FP16 weights and activations, FP32
// accumulation); hidden units are the M dimension and states the N
// dimension, so a layer's accumulator registers are already the next layer's
// B operand.  The plain solve stays the judge: the Newton iteration and its
// tolerance are unchanged, and a state whose Newton run from the learned
// guess fails is solved again from the A27 guess.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "physics.hpp"

namespace gpl {

// weight vector layout (floats)
constexpr int NET_H = 32;
constexpr int NET_MU = 0, NET_SD = 4, NET_W1 = 8, NET_B1 = NET_W1 + NET_H*4;
constexpr int NET_W2 = NET_B1 + NET_H, NET_B2 = NET_W2 + NET_H*NET_H;
constexpr int NET_W3 = NET_B2 + NET_H, NET_B3 = NET_W3 + NET_H, NET_YSC = NET_B3 + 1;
constexpr int NET_SIZE = NET_YSC + 2;          // 1259; the last float is reserved
constexpr int NET_TRAIN_BEG = NET_W1, NET_TRAIN_END = NET_B3 + 1;   // trainable range

// Scale-free invariants of u = (D, m1..3, B1..3, E - D), raw (not normalised).
__device__ __forceinline__ void net_features(const double *u, float *x)
{
  const double D  = u[0];
  const double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  const double B2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
  const double S  = u[1]*u[4] + u[2]*u[5] + u[3]*u[6];
  x[0] = (float)log(u[7]/D);
  x[1] = (float)log1p(sqrt(m2)/D);
  x[2] = (float)log1p(B2/D);
  x[3] = (float)(S*S/(m2*B2 + 1e-300));
#pragma unroll
  for (int q = 0; q < 4; q++) if (!(x[q] == x[q]) || fabsf(x[q]) > 1e30f) x[q] = 0.f;
}

typedef _Float16 net_half4 __attribute__((ext_vector_type(4)));
typedef float net_float16 __attribute__((ext_vector_type(16)));

// One network pass over 32 states on the matrix cores.  Every lane of the
// wavefront must call it.  Lane l < 32 supplies the raw features of state l
// in xs[0..3] (lanes 32..63 pass anything); on return lanes l and l + 32
// both hold the network output of state l % 32.
__device__ __forceinline__ float net_eval32(const float *__restrict__ net, const float *xs)
{
  const int lane = threadIdx.x & 63, col = lane & 31, hi = lane >> 5;
  // layer 1: A = W1 (hidden x 8 padded features), B = x^T (8 x 32 states)
  net_half4 a, b;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int f = 4*hi + q;
    a[q] = (_Float16)(f < 4 ? net[NET_W1 + col*4 + f] : 0.f);
    b[q] = (_Float16)(hi == 0 ? (xs[q] - net[NET_MU + q])/net[NET_SD + q] : 0.f);
  }
  net_float16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.f;
  acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, acc, 0, 0, 0);
  // accumulator register r of lane l holds hidden unit 8*(r/4) + 4*hi + r%4 of state col
  float h[16];
#pragma unroll
  for (int r = 0; r < 16; r++) h[r] = tanhf(acc[r] + net[NET_B1 + 8*(r >> 2) + 4*hi + (r & 3)]);
  // layer 2: four K-steps of 8 hidden units; B operand = registers 4s..4s+3
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.f;
#pragma unroll
  for (int s = 0; s < 4; s++) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      a[q] = (_Float16)net[NET_W2 + col*NET_H + 8*s + 4*hi + q];
      b[q] = (_Float16)h[4*s + q];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, acc, 0, 0, 0);
  }
  // layer 3 (VALU): the lane's 16 hidden units, then the other half-wave's
  float o = 0.f;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int u = 8*(r >> 2) + 4*hi + (r & 3);
    o += net[NET_W3 + u]*tanhf(acc[r] + net[NET_B2 + u]);
  }
  o += __shfl_xor(o, 32, 64);
  return o + net[NET_B3];
}

// Learned guess for the state held by this lane (all 64 lanes call; 'xs' are
// this lane's raw features).  Two passes: states of lanes 0..31, then 32..63.
__device__ __forceinline__ float net_eval64(const float *__restrict__ net, const float *x)
{
  const int lane = threadIdx.x & 63;
  float xs[4];
#pragma unroll
  for (int q = 0; q < 4; q++) xs[q] = x[q];
  const float o1 = net_eval32(net, xs);
#pragma unroll
  for (int q = 0; q < 4; q++) xs[q] = __shfl_xor(x[q], 32, 64);   // lane l < 32 gets state l + 32
  const float o2 = net_eval32(net, xs);
  return lane < 32 ? o1 : o2;
}

// Newton from the learned guess, the A27 guess as the fallback and judge.
// Returns the solve's error bits (0 on success); *passes counts every pass.
__device__ __forceinline__ uint32_t rmhd_energy_solve_guided(const double *u, double *v, double gamma,
                                                             float pred, float ysc, int *passes,
                                                             int *restart = nullptr)
{
  int it1 = 0, it2 = 0;
  const double wa = rmhd_guess_a27(u);
  const double w0 = wa*exp((double)(ysc*pred));
  uint32_t e = rmhd_energy_solve(u, v, gamma, &it1, w0);
  const bool again = e != 0;
  if (again) e = rmhd_energy_solve(u, v, gamma, &it2);
  if (passes) *passes = it1 + it2;
  if (restart) *restart = again ? 1 : 0;
  return e;
}

}  // namespace gpl
