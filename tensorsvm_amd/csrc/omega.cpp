// Random test matrix Omega: the reference's gaussrand() stream, bit-identical and
// multithreaded.
//
// tensorsvm-train.cu:2621-2643 draws N(0,1) values with the Marsaglia polar method from
// glibc rand(), which the reference never seeds (so it starts from glibc's seed-1
// state), and fills Omega row-major (1966-1974).  glibc's rand() is the additive
// lagged-Fibonacci generator of random_r (TYPE_3, degree 31, separation 3):
//   r_0 = seed, r_i = 16807 r_{i-1} mod (2^31 - 1) for i < 31, r_i = r_{i-31} for
//   31 <= i < 34, r_i = r_{i-31} + r_{i-3} (mod 2^32) afterwards; output k is
//   r_{k+344} >> 1.
// The recurrence is linear over Z/2^32, so the state can jump ahead by a matrix power;
// the raw stream is split into blocks of pairs, each thread counts the accepted pairs
// of its blocks, a prefix sum fixes where each block's outputs go, and a second pass
// writes them.  Float arithmetic follows the reference exactly (float log/sqrt
// overloads = logf/sqrtf; no contraction: this file is built with -ffp-contract=off).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/tsvm.h"

namespace {

constexpr int DEG = 31;
typedef uint32_t Mat[DEG][DEG];

struct State {
  uint32_t s[DEG];  // s[q] = r_{i-31+q}; the next value is s[0] + s[28]
};

State initial_state() {
  int64_t r[400];
  r[0] = 1;
  for (int i = 1; i < 31; i++) {
    int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
    int64_t w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    r[i] = w;
  }
  uint32_t u[400];
  for (int i = 0; i < 31; i++) u[i] = (uint32_t)r[i];
  for (int i = 31; i < 34; i++) u[i] = u[i - 31];
  for (int i = 34; i < 344; i++) u[i] = u[i - 31] + u[i - 3];
  State st;
  for (int q = 0; q < DEG; q++) st.s[q] = u[344 - 31 + q];
  return st;
}

void mat_mul(const Mat& A, const Mat& B, Mat& C) {
  Mat T;
  for (int i = 0; i < DEG; i++)
    for (int j = 0; j < DEG; j++) {
      uint32_t acc = 0;
      for (int t = 0; t < DEG; t++) acc += A[i][t] * B[t][j];
      T[i][j] = acc;
    }
  memcpy(C, T, sizeof(Mat));
}

// J = T^steps where T advances the state by one raw value
void jump_matrix(uint64_t steps, Mat& J) {
  Mat T, R;
  memset(T, 0, sizeof(T));
  for (int q = 0; q < DEG - 1; q++) T[q][q + 1] = 1;
  T[DEG - 1][0] = 1;
  T[DEG - 1][28] = 1;
  memset(R, 0, sizeof(R));
  for (int q = 0; q < DEG; q++) R[q][q] = 1;
  while (steps) {
    if (steps & 1) mat_mul(R, T, R);
    mat_mul(T, T, T);
    steps >>= 1;
  }
  memcpy(J, R, sizeof(Mat));
}

State apply(const Mat& J, const State& a) {
  State b;
  for (int i = 0; i < DEG; i++) {
    uint32_t acc = 0;
    for (int t = 0; t < DEG; t++) acc += J[i][t] * a.s[t];
    b.s[i] = acc;
  }
  return b;
}

// Raw generator over a 31-word ring.
struct Gen {
  uint32_t ring[DEG];
  int head = 0;  // ring[head] = r_{i-31}
  explicit Gen(const State& st) { memcpy(ring, st.s, sizeof(ring)); }
  inline int32_t next() {
    int i3 = head + 28;
    if (i3 >= DEG) i3 -= DEG;
    uint32_t v = ring[head] + ring[i3];
    ring[head] = v;
    head = head + 1 == DEG ? 0 : head + 1;
    return (int32_t)(v >> 1);
  }
};

// one polar-method attempt on a raw pair; returns whether accepted and the two outputs
inline bool polar(int32_t r1, int32_t r2, float* o1, float* o2) {
  const float rmax = 2147483647.0f;  // RAND_MAX converted to float (= 2^31)
  float U1 = (float)r1 / rmax;
  float U2 = (float)r2 / rmax;
  float V1 = 2.0f * U1 - 1.0f;
  float V2 = 2.0f * U2 - 1.0f;
  float a = V1 * V1;
  float b = V2 * V2;
  float S = a + b;
  if (S >= 1.0f || S == 0.0f) return false;
  float l = logf(S);
  float q = (-2.0f * l) / S;
  float f = sqrtf(q);
  *o1 = V1 * f;
  *o2 = V2 * f;
  return true;
}

}  // namespace

extern "C" tsvm_status tsvm_gaussrand_fill(float* omega, int64_t count, int nthreads) {
  if (count < 0 || (count > 0 && !omega)) return TSVM_ERR_ARG;
  if (count == 0) return TSVM_OK;
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  const int64_t BP = int64_t(1) << 18;  // raw pairs per block
  const int64_t pairs_needed = (count + 1) / 2;
  // about 78.5% of pairs are accepted; overshoot and extend if needed
  int64_t nblocks = pairs_needed * 4 / 3 / BP + 2;
  Mat J;
  jump_matrix((uint64_t)(2 * BP), J);
  std::vector<State> st;
  std::vector<int64_t> acc;
  st.push_back(initial_state());
  for (;;) {
    while ((int64_t)st.size() < nblocks) st.push_back(apply(J, st.back()));
    acc.assign(nblocks, 0);
    // pass 1: accepted pairs per block
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
      th.emplace_back([&, t]() {
        for (int64_t b = t; b < nblocks; b += nthreads) {
          Gen g(st[b]);
          int64_t a = 0;
          for (int64_t p = 0; p < BP; p++) {
            int32_t r1 = g.next(), r2 = g.next();
            float o1, o2;
            if (polar(r1, r2, &o1, &o2)) a++;
          }
          acc[b] = a;
        }
      });
    for (auto& x : th) x.join();
    int64_t tot = 0;
    for (int64_t b = 0; b < nblocks; b++) tot += acc[b];
    if (tot >= pairs_needed) break;
    nblocks += nblocks / 4 + 2;
  }
  // prefix: first output pair index of each block
  std::vector<int64_t> off(nblocks + 1, 0);
  for (int64_t b = 0; b < nblocks; b++) off[b + 1] = off[b] + acc[b];
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([&, t]() {
      for (int64_t b = t; b < nblocks; b += nthreads) {
        int64_t pidx = off[b];
        if (pidx >= pairs_needed) continue;
        Gen g(st[b]);
        for (int64_t p = 0; p < BP && pidx < pairs_needed; p++) {
          int32_t r1 = g.next(), r2 = g.next();
          float o1, o2;
          if (!polar(r1, r2, &o1, &o2)) continue;
          int64_t o = 2 * pidx;
          omega[o] = o1;
          if (o + 1 < count) omega[o + 1] = o2;
          pidx++;
        }
      }
    });
  for (auto& x : th) x.join();
  return TSVM_OK;
}
