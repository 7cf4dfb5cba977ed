// Host build of the device's correctly rounded acos/sincos (csrc/crmath.hpp),
// compared with the host libm (glibc) that the reference's CPU build calls.
// Random arguments come from a counter-based generator (splitmix64 of the
// index), so every count is reproducible; the work is split over threads.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../gpluto_dev_amd/csrc/crmath.hpp"

namespace {

uint64_t splitmix(uint64_t x)
{
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30))*0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27))*0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
double unit(uint64_t i, uint64_t seed) { return (splitmix(i ^ (seed*0x632be59bd9b4e019ull)) >> 11)*0x1p-53; }

// mode 0: acos on uniform [lo, hi]; 1: acos with x = +-(1 - 2^-e u), e in
// [1, 52] (arguments that crowd at +-1); 2: sincos on uniform [lo, hi];
// 3: sincos of theta = acos(arg)/3 for uniform arg in [-1, 1] (the cubic's
// own theta); 4: acos_refine from a seed 1..4 ulp off the library's, against
// acos_cr (not glibc).
double argument(int mode, uint64_t i, uint64_t seed, double lo, double hi)
{
  const double u = unit(i, seed);
  if (mode == 1) {
    const int e = 1 + (int)(splitmix(i + 0x5bd1e995ull*seed) % 52);
    const double x = 1.0 - ldexp(u, -e);
    return (splitmix(i*3 + seed) & 1) ? x : -x;
  }
  if (mode == 3) return acos(-1.0 + 2.0*u)*(1.0/3.0);
  return lo + (hi - lo)*u;
}

bool mismatch(int mode, uint64_t i, uint64_t seed, double x)
{
  if (mode == 2 || mode == 3) {
    double s, c, s0, c0;
    gpl::cr::sincos_cr(x, &s, &c);
    sincos(x, &s0, &c0);
    return s != s0 || c != c0;
  }
  const double ref = acos(x);
  if (mode == 4) {
    if (x == 1.0 || x == -1.0) return false;
    double seed_v = gpl::cr::acos_seed(x);
    const int k = 1 + (int)(splitmix(i ^ 0xabcdefull) % 4);
    const double dir = (splitmix(i + seed) & 1) ? INFINITY : -INFINITY;
    for (int j = 0; j < k; j++) seed_v = nextafter(seed_v, dir);
    return gpl::cr::acos_refine(x, seed_v) != gpl::cr::acos_cr(x);
  }
  return gpl::cr::acos_cr(x) != ref;
}

}  // namespace

extern "C" {

double crm_acos(double x) { return gpl::cr::acos_cr(x); }
void crm_sincos(double x, double *s, double *c) { gpl::cr::sincos_cr(x, s, c); }

// Count the arguments i in [0, n) where the correctly rounded routine and
// glibc disagree; the first few mismatching arguments go to first[nfirst].
long long crm_compare(int mode, long long n, unsigned long long seed, double lo, double hi,
                      int nthreads, double *first, int nfirst)
{
  std::atomic<long long> total{0};
  std::atomic<int> stored{0};
  std::vector<std::thread> pool;
  nthreads = std::max(1, nthreads);
  for (int t = 0; t < nthreads; t++) {
    pool.emplace_back([&, t]() {
      long long cnt = 0;
      for (long long i = t; i < n; i += nthreads) {
        const double x = argument(mode, (uint64_t)i, seed, lo, hi);
        if (mismatch(mode, (uint64_t)i, seed, x)) {
          cnt++;
          const int slot = stored.fetch_add(1);
          if (slot < nfirst) first[slot] = x;
        }
      }
      total += cnt;
    });
  }
  for (auto &th : pool) th.join();
  return total.load();
}

}  // extern "C"
