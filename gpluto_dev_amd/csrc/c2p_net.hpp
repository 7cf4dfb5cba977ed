// Learned initial guess for the RMHD con2prim Newton solve.
//
// An 8-64-64-1 ReLU network maps eight scale-free invariants of the conserved
// state to a correction of the Eq. (A27) guess of RMHD_EnergySolve
// (Src/RMHD/rmhd_energy_solve.cpp:86-96): W_guess = W_A27 * exp(ysc * net(x)).
//
// Inference runs on the matrix cores with v_mfma_f32_32x32x16_bf16 (BF16
// weights and activations, FP32 accumulation).  Hidden units are the M
// dimension and states the N dimension, so a layer's accumulator registers
// are, after ReLU and conversion, the next layer's B operand in the same
// lane.  A wavefront evaluates its 64 states in two passes of 32; the pass
// that serves lanes 32..63 swaps the two K halves of the first layer and
// puts the output in row 4 instead of row 0, so no value crosses lanes.
// Biases ride on a constant-1 input slot (layer 1) or a constant-1 K block
// (layer 2 and the output), so the accumulators start at zero.
//
// The plain solve stays the arbiter: the Newton iteration and its tolerance
// are unchanged, and a state whose Newton run from the learned guess fails
// is solved again from the A27 guess.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "physics.hpp"

namespace gpl {

// the float weight vector (the ABI's form)
constexpr int NET_F = 8, NET_H = 64;
constexpr int NET_MU = 0, NET_SD = NET_F, NET_W1 = 2*NET_F, NET_B1 = NET_W1 + NET_H*NET_F;
constexpr int NET_W2 = NET_B1 + NET_H, NET_B2 = NET_W2 + NET_H*NET_H;
constexpr int NET_W3 = NET_B2 + NET_H, NET_B3 = NET_W3 + NET_H, NET_YSC = NET_B3 + 1;
constexpr int NET_SIZE = NET_YSC + 2;          // 4819; the last float is reserved
constexpr int NET_TRAIN_BEG = NET_W1, NET_TRAIN_END = NET_B3 + 1;   // trainable range

// The matrix-core form, built on the device from the weight vector
// (k_net_pack): BF16 A operands in the order each lane loads them (8 per
// lane and MFMA, one 16-byte load), and the FP32 input normalisation.
//   l1[pass][mtile][lane][8]  layer 1: W1 rows, b1 on the constant slot
//   l2[mtile][kblock][lane][8] layer 2: W2 in the K order of layer 1's
//                              accumulator registers; kblock 4 carries b2
//   l3[pass][kblock][lane][8] output: W3 (+ b3) in row 0 (pass 0) or row 4
struct NetPack {
  uint16_t l1[2][2][64][8];
  uint16_t l2[2][5][64][8];
  uint16_t l3[2][5][64][8];
  float mu[NET_F], rsd[NET_F], ysc, pad[15];
};
static_assert(sizeof(NetPack) % 64 == 0, "NetPack keeps 64-byte alignment");

// Weight vector (host) -> NetPack (device), through the device copy dev_w
// of the vector, on stream s (c2p_net.hip).
int net_pack_upload(const float *weights, float *dev_w, NetPack *dev_p, hipStream_t s);

// hidden unit carried by K slot 8*hi + j of K block s (s < 4): the layout of
// the 32x32 accumulator, whose register r of lane l holds row
// 8*(r/4) + 4*(l/32) + r%4 of its M tile
__host__ __device__ constexpr int net_unit(int s, int hi, int j)
{
  return 16*s + 8*(j/4) + 4*hi + (j%4);
}

// Scale-free invariants of u = (D, m1..3, B1..3, tau = E - D), raw (not
// normalised), in FP32; wa = the A27 root.  Non-finite values read as 0.
__device__ __forceinline__ void net_features(const double *u, double wa, float *x)
{
  const float D = (float)u[0], tau = (float)u[7];
  const float m1 = (float)u[1], m2c = (float)u[2], m3 = (float)u[3];
  const float b1 = (float)u[4], b2c = (float)u[5], b3 = (float)u[6];
  const float mm = m1*m1 + m2c*m2c + m3*m3;
  const float bb = b1*b1 + b2c*b2c + b3*b3;
  const float s  = m1*b1 + m2c*b2c + m3*b3;
  const float et = tau + D;
  const float ret = __builtin_amdgcn_rcpf(et), rd = __builtin_amdgcn_rcpf(D);
  const float mabs = __builtin_amdgcn_sqrtf(mm);
  x[0] = __logf(D*ret);
  x[1] = mabs*ret;
  x[2] = bb*ret;
  x[3] = s*s*__builtin_amdgcn_rcpf(mm*bb + 1e-30f);
  x[4] = __logf(1.f + bb*rd);
  x[5] = __logf(1.f + mabs*rd);
  x[6] = __logf(fmaxf(tau*rd, 1e-30f));
  x[7] = (float)wa*ret;
#pragma unroll
  for (int q = 0; q < NET_F; q++) if (!(x[q] == x[q]) || fabsf(x[q]) > 1e30f) x[q] = 0.f;
}

typedef __bf16 net_bf8 __attribute__((ext_vector_type(8)));
typedef short net_s8 __attribute__((ext_vector_type(8)));
typedef float net_f8 __attribute__((ext_vector_type(8)));
typedef float net_f16 __attribute__((ext_vector_type(16)));
typedef unsigned int net_u4 __attribute__((ext_vector_type(4)));

typedef float net_f2 __attribute__((ext_vector_type(2)));
typedef __bf16 net_bf2 __attribute__((ext_vector_type(2)));
typedef short net_s2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ net_bf8 net_ld8(const uint16_t *p)
{
  return __builtin_bit_cast(net_bf8, *(const net_u4 *)p);
}

// Two floats as a BF16 pair (round to nearest even): one v_cvt_pk_bf16_f32.
__device__ __forceinline__ unsigned int net_cvt2(float x, float y)
{
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(net_f2{x, y}, net_bf2));
}

// ReLU of a BF16 pair: max with +0 on the bit patterns as int16 (a BF16 with
// its sign bit set is a negative int16).
__device__ __forceinline__ unsigned int net_relu2(unsigned int p)
{
  return __builtin_bit_cast(unsigned int, __builtin_elementwise_max(__builtin_bit_cast(net_s2, p), (net_s2)0));
}

// (written out pair by pair: a loop assigning the elements of a vector
// variable was folded into a splat of its first element by this compiler)
__device__ __forceinline__ net_bf8 net_cvt8(net_f8 a)
{
  const net_u4 r = {net_cvt2(a[0], a[1]), net_cvt2(a[2], a[3]), net_cvt2(a[4], a[5]),
                    net_cvt2(a[6], a[7])};
  return __builtin_bit_cast(net_bf8, r);
}

__device__ __forceinline__ net_bf8 net_relu8(net_f8 a)
{
  const net_u4 r = {net_relu2(net_cvt2(a[0], a[1])), net_relu2(net_cvt2(a[2], a[3])),
                    net_relu2(net_cvt2(a[4], a[5])), net_relu2(net_cvt2(a[6], a[7]))};
  return __builtin_bit_cast(net_bf8, r);
}

__device__ __forceinline__ net_f8 net_take8(const net_f16 &acc, int half)
{
  net_f8 r;
#pragma unroll
  for (int j = 0; j < 8; j++) r[j] = acc[8*half + j];
  return r;
}

// One pass over 32 states: pass 0 serves the states of lanes 0..31, pass 1
// those of lanes 32..63.  'xb' is this lane's normalised features as BF16;
// returns register 0 of the output tile (the prediction for the lane's own
// state in the lanes the pass serves).
__device__ __forceinline__ float net_pass(const NetPack *__restrict__ P, int pass, net_bf8 xb)
{
  const int lane = threadIdx.x & 63, hi = lane >> 5;
  net_bf8 one;
#pragma unroll
  for (int j = 0; j < 8; j++) one[j] = (__bf16)(j == 0 ? 1.f : 0.f);
  const net_bf8 zero8 = __builtin_bit_cast(net_bf8, (net_s8)0);
  // layer 1: the serving half supplies the features, the other half the
  // constant slot that carries b1
  const net_bf8 b1op = hi == pass ? xb : one;
  net_f16 h1[2];
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
    net_f16 acc = {};
    h1[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(net_ld8(P->l1[pass][mt][lane]), b1op, acc, 0, 0, 0);
  }
  net_bf8 a1[4];
#pragma unroll
  for (int s = 0; s < 4; s++) a1[s] = net_relu8(net_take8(h1[s >> 1], s & 1));
  const net_bf8 bias_op = hi == 0 ? one : zero8;
  // layer 2
  net_f16 h2[2];
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
    net_f16 acc = {};
#pragma unroll
    for (int s = 0; s < 4; s++)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(net_ld8(P->l2[mt][s][lane]), a1[s], acc, 0, 0, 0);
    h2[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(net_ld8(P->l2[mt][4][lane]), bias_op, acc, 0, 0, 0);
  }
  // output layer: one useful row
  net_f16 o = {};
#pragma unroll
  for (int s = 0; s < 4; s++)
    o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(net_ld8(P->l3[pass][s][lane]),
                                                net_relu8(net_take8(h2[s >> 1], s & 1)), o, 0, 0, 0);
  o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(net_ld8(P->l3[pass][4][lane]), bias_op, o, 0, 0, 0);
  return o[0];
}

// Learned correction ysc * net(x) for the state held by this lane.  Every lane
// of the wavefront must call it; 'x' is this lane's raw features.
__device__ __forceinline__ float net_eval64(const NetPack *__restrict__ P, const float *x)
{
  const int lane = threadIdx.x & 63;
  net_f8 xn;
#pragma unroll
  for (int q = 0; q < NET_F; q++) xn[q] = (x[q] - P->mu[q])*P->rsd[q];
  const net_bf8 xb = net_cvt8(xn);
  const float o0 = net_pass(P, 0, xb);
  const float o1 = net_pass(P, 1, xb);
  return P->ysc*(lane < 32 ? o0 : o1);
}

// Newton from the learned guess, the A27 guess as the fallback.
// Returns the solve's error bits (0 on success); *passes counts every pass.
__device__ __forceinline__ uint32_t rmhd_energy_solve_guided(const double *u, double *v, double gamma,
                                                             double wa, float corr, int *passes,
                                                             int *restart = nullptr)
{
  int it1 = 0, it2 = 0;
  const double w0 = wa*(double)__expf(corr);
  uint32_t e = rmhd_energy_solve(u, v, gamma, &it1, w0);
  const bool again = e != 0;
  if (again) e = rmhd_energy_solve(u, v, gamma, &it2);
  if (passes) *passes = it1 + it2;
  if (restart) *restart = again ? 1 : 0;
  return e;
}

}  // namespace gpl
