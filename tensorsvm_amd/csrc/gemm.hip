// Tall-skinny dense GEMMs of the factor construction, on FP64 MFMA
// (v_mfma_f64_16x16x4f64, FP64 accumulation of FP32 operands):
//
//   gemm_tn_f64:  C (ka x kb) = A^T diag(w) B over n rows   (Gram W^T W, C = W^T Q,
//                 the plain-form capacitance Z^T D^-1 Z)   -- replaces the chunked
//                 cublasDgemm of tensorsvm-train.cu:2122-2149 and cblas_dsyrk/cublasDsyrk
//                 of 1674-1697
//   gemm_nn_tall: Y (n x kb, FP32) = A (n x ka, FP32) * B (ka x kb, FP64)   (W R^-1,
//                 U = W X) -- replaces cublasSgemm of 799-806 and 2231-2233
//
// gemm_tn_f64 splits the n rows over a third grid dimension; each split writes its
// partial tile and a second kernel sums the splits in a fixed order (deterministic).
#include <type_traits>

#include "common.h"
#include "internal.h"

namespace tsvm {

// 256 threads; the block owns a TILE x TILE output tile (TILE = 64 or 128), wave w a
// TILE/2 x TILE/2 quarter (TILE/32 x TILE/32 MFMA tiles of 16 x 16; the 128 form has 16
// independent accumulators per wave and reads half an LDS fragment per MFMA).  Rows are
// staged TN_R at a time through LDS as FP64 (the A panel already multiplied by the row
// weight), double-buffered in LDS, with the global loads two stages ahead in a register
// ring.  LDS rows are TILE doubles (a multiple of the bank count), so odd rows store
// their columns XOR 16: the four row groups of a fragment read (lanes 16 q .. 16 q + 15
// read row 4 s + q) then fall on disjoint banks within each half-wave instead of all on
// the same 32.
constexpr int TN_R = 16;  // rows per stage
constexpr int TN_H = TN_R / 16;  // staging passes of 16 rows
__device__ inline int tn_sw(int row, int col) { return col ^ ((row & 1) << 4); }

// IDX: the n rows taken are rows[0..n) of A, B and w (an index list; the index of a stage's
// row is loaded one stage before its data, so the data loads do not wait on it).
template <bool VEC, bool HAS_W, bool IDX = false, int TILE = 64>
__global__ __launch_bounds__(256) void gemm_tn_f64_kernel(
    const float* __restrict__ A, int64_t lda, int64_t ka, const float* __restrict__ B,
    int64_t ldb, int64_t kb, int64_t n, const double* __restrict__ w, int lower_only,
    int64_t rows_per_split, int64_t nsplit, double* __restrict__ part,
    const int* __restrict__ rows = nullptr) {
  constexpr int CG = TILE / 64;  // 64-column groups of a staged row
  constexpr int NI = TILE / 32;  // MFMA tiles per wave along each side
  // 1-D grid, XCD-aware: blocks go to the 8 XCDs round-robin by index, and an XCD's
  // consecutive blocks take the tiles of one row split in turn, so the split's rows are
  // read from HBM once and from the XCD's L2 for its other tiles.  lower_only: tile t
  // enumerates the lower tiles row by row, (ta, tb) with ta (ta + 1) / 2 + tb = t,
  // tb <= ta (no upper tiles in the grid: idle ones would unbalance the XCDs).
  const int nta = (int)((ka + TILE - 1) / TILE), ntb = (int)((kb + TILE - 1) / TILE);
  const int64_t nt = lower_only ? (int64_t)nta * (nta + 1) / 2 : (int64_t)nta * ntb;
  const int64_t bid = blockIdx.x, jx = bid / 8;
  const int64_t split = (jx / nt) * 8 + bid % 8;
  if (split >= nsplit) return;
  const int t = (int)(jx % nt);
  int ta = t % nta, tb = t / nta;
  if (lower_only) {
    ta = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((ta + 1) * (ta + 2) / 2 <= t) ta++;
    while (ta * (ta + 1) / 2 > t) ta--;
    tb = t - ta * (ta + 1) / 2;
  }
  __shared__ double sA[2][TN_R][TILE];
  __shared__ double sB[2][TN_R][TILE];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int wa = (wave >> 1) * (TILE / 2), wb = (wave & 1) * (TILE / 2);
  const int64_t a0 = (int64_t)ta * TILE, b0 = (int64_t)tb * TILE;
  const int64_t r0 = split * rows_per_split;
  const int64_t r1 = r0 + rows_per_split < n ? r0 + rows_per_split : n;
  const int64_t nst = (r1 - r0 + TN_R - 1) / TN_R;
  // staging map: thread -> (row, 4 columns in each 64-column group) for TN_H passes
  const int sr = tid >> 4, sc4 = (tid & 15) * 4;
  f32x4_t ra[2][TN_H][CG], rb[2][TN_H][CG];  // [ring slot][h][column group]
  double rw[2][TN_H];
  bool rok[2][TN_H];
  // rows past the split load a valid row and are zeroed when staged: no branches, and
  // nothing waits on the loads before the data is needed.  With IDX, pidx holds the
  // physical rows of the stage being loaded (fetched one stage earlier by iload).
  int pidx[TN_H];
  auto iload = [&](int64_t st) {
#pragma unroll
    for (int h = 0; h < TN_H; h++) {
      const int64_t row = r0 + st * TN_R + sr + 16 * h;
      pidx[h] = IDX ? rows[row < r1 ? row : r0] : 0;
    }
  };
  auto gload = [&](int64_t st, int slot) {
#pragma unroll
    for (int h = 0; h < TN_H; h++) {
      const int64_t row = r0 + st * TN_R + sr + 16 * h;
      const bool ok = row < r1;
      const int64_t rr = IDX ? (int64_t)pidx[h] : (ok ? row : r0);
      rok[slot][h] = ok;
      rw[slot][h] = HAS_W ? w[rr] : 1.0;
#pragma unroll
      for (int cg = 0; cg < CG; cg++) {
        const int64_t ca0 = a0 + cg * 64 + sc4, cb0 = b0 + cg * 64 + sc4;
        f32x4_t va, vb;
        if (VEC) {
          va = *(const f32x4_t*)(A + rr * lda + ca0);
          vb = *(const f32x4_t*)(B + rr * ldb + cb0);
        } else {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const bool ca = ca0 + j < ka, cb = cb0 + j < kb;
            const float xa = A[rr * lda + (ca ? ca0 + j : 0)];
            const float xb = B[rr * ldb + (cb ? cb0 + j : 0)];
            va[j] = ca ? xa : 0.0f;
            vb[j] = cb ? xb : 0.0f;
          }
        }
        ra[slot][h][cg] = va;
        rb[slot][h][cg] = vb;
      }
    }
  };
  auto sstore = [&](int buf, int slot) {
#pragma unroll
    for (int h = 0; h < TN_H; h++) {
      const int r = sr + 16 * h;
      const double wv = rok[slot][h] ? rw[slot][h] : 0.0;
#pragma unroll
      for (int cg = 0; cg < CG; cg++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          sA[buf][r][tn_sw(r, cg * 64 + sc4 + j)] = wv * (double)ra[slot][h][cg][j];
          sB[buf][r][tn_sw(r, cg * 64 + sc4 + j)] = rok[slot][h] ? (double)rb[slot][h][cg][j] : 0.0;
        }
    }
  };
  f64x4_t acc[NI][NI];
#pragma unroll
  for (int i = 0; i < NI; i++)
#pragma unroll
    for (int j = 0; j < NI; j++) acc[i][j] = (f64x4_t){0, 0, 0, 0};
  auto compute = [&](int buf) {
#pragma unroll
    for (int s4 = 0; s4 < TN_R / 4; s4++) {
      const int row = 4 * s4 + kq;
      double av[NI], bv[NI];
#pragma unroll
      for (int i = 0; i < NI; i++) {
        av[i] = sA[buf][row][tn_sw(row, wa + 16 * i + l16)];
        bv[i] = sB[buf][row][tn_sw(row, wb + 16 * i + l16)];
      }
#pragma unroll
      for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < NI; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  };
  // One stage with its ring slot and LDS buffer fixed at compile time (b): register slot
  // b held stage st, already in LDS buffer b, and is refilled with stage st + 2; after
  // the MFMAs stage st + 1 (slot 1 - b) goes into LDS buffer 1 - b, last read before the
  // previous barrier.  Stages past the split stage zeros, so stages run in unconditional
  // pairs (a conditional stage would move the accumulators between register files).
  auto stage = [&](int64_t st, auto bc) {
    constexpr int b = decltype(bc)::value;
    gload(st + 2, b);
    if (IDX) iload(st + 3);
    __builtin_amdgcn_sched_barrier(0);
    compute(b);
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    sstore(1 - b, 1 - b);
    __syncthreads();
  };
  if (nst > 0) {
    if (IDX) iload(0);
    gload(0, 0);
    if (IDX) iload(1);
    gload(1, 1);
    if (IDX) iload(2);
    sstore(0, 0);
    __syncthreads();
  }
  for (int64_t st = 0; st < nst; st += 2) {
    stage(st, std::integral_constant<int, 0>());
    stage(st + 1, std::integral_constant<int, 1>());
  }
  // D layout of v_mfma_f64_16x16x4f64: item q of a lane holds row lane/16 + 4q,
  // column lane%16
  double* P = part + split * ka * kb;
#pragma unroll
  for (int i = 0; i < NI; i++)
#pragma unroll
    for (int j = 0; j < NI; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int64_t a = a0 + wa + 16 * i + kq + 4 * q, b = b0 + wb + 16 * j + l16;
        if (a < ka && b < kb) P[a * kb + b] = acc[i][j][q];
      }
}

__global__ void reduce_splits_kernel(const double* __restrict__ part, int64_t nsplit,
                                     int64_t ka, int64_t kb, int lower_only,
                                     double* __restrict__ C) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= ka * kb) return;
  int64_t a = idx / kb, b = idx - a * kb;
  // lower_only: the lower triangle is summed and mirrored entry by entry, so C is
  // exactly symmetric (inside a diagonal tile (w z_a) z_b and (w z_b) z_a round apart)
  if (lower_only && b > a) return;
  double s = 0.0;
  for (int64_t t = 0; t < nsplit; t++) s += part[t * ka * kb + idx];
  C[idx] = s;
  if (lower_only && b < a) C[b * kb + a] = s;
}

tsvm_status gemm_tn_f64(Ctx* c, const float* A, int64_t lda, int64_t ka, const float* B,
                        int64_t ldb, int64_t kb, int64_t n, const double* w, bool lower_only,
                        double* C, const int* rows) {
  if (ka <= 0 || kb <= 0) return TSVM_OK;
  if (lower_only && ka != kb) return fail(c, TSVM_ERR_ARG, "gemm_tn_f64: lower_only needs ka==kb");
  if (rows && (A != B || lda != ldb || ka != kb))
    return fail(c, TSVM_ERR_ARG, "gemm_tn_f64: a row list needs A == B");
  // 128-column tiles for wide operands (16 accumulators per wave), 64 otherwise
  const int tile = ka >= 256 && kb >= 256 ? 128 : 64;
  const int64_t ta = ceil_div(ka, tile), tb = ceil_div(kb, tile);
  const int64_t tiles = lower_only ? ta * (ta + 1) / 2 : ta * tb;
  // enough blocks to fill the chip several times; at least 256 rows per split
  int64_t nsplit = ceil_div((int64_t)c->cu_count * 8, tiles);
  int64_t max_split = ceil_div(n > 0 ? n : 1, 256);
  if (nsplit > max_split) nsplit = max_split;
  if (nsplit < 1) nsplit = 1;
  int64_t rps = round_up(ceil_div(n > 0 ? n : 1, nsplit), TN_R);
  nsplit = ceil_div(n > 0 ? n : 1, rps);
  void* part;
  TSVM_HIP(c, c->ws.get("gemm_part", (size_t)(nsplit * ka * kb * 8), &part));
  // whole rounds of the 8 XCDs (the splits past nsplit return at once)
  const dim3 grid((unsigned)(round_up(nsplit, 8) * tiles));
  // whole tiles with 16-byte rows: vector loads without column masks
  const bool vec = ka % tile == 0 && kb % tile == 0 && lda % 4 == 0 && ldb % 4 == 0;
#define GTN(V, W, I)                                                                              \
  do {                                                                                          \
    if (tile == 128)                                                                            \
      hipLaunchKernelGGL((gemm_tn_f64_kernel<V, W, I, 128>), grid, dim3(256), 0, c->stream, A,   \
                         lda, ka, B, ldb, kb, n, w, lower_only ? 1 : 0, rps, nsplit,             \
                         (double*)part, rows);                                                  \
    else                                                                                        \
      hipLaunchKernelGGL((gemm_tn_f64_kernel<V, W, I, 64>), grid, dim3(256), 0, c->stream, A,    \
                         lda, ka, B, ldb, kb, n, w, lower_only ? 1 : 0, rps, nsplit,             \
                         (double*)part, rows);                                                  \
  } while (0)
  if (rows) {
    // the FP64 rows of the hybrid capacitance: rows read in place through the index list
    if (vec) {
      if (w) GTN(true, true, true); else GTN(true, false, true);
    } else {
      if (w) GTN(false, true, true); else GTN(false, false, true);
    }
  } else if (vec) {
    if (w) GTN(true, true, false); else GTN(true, false, false);
  } else {
    if (w) GTN(false, true, false); else GTN(false, false, false);
  }
#undef GTN
  hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)ceil_div(ka * kb, 256)), dim3(256), 0,
                     c->stream, (const double*)part, nsplit, ka, kb, lower_only ? 1 : 0, C);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

// Generic form: 256 threads; block owns 64 rows x 64 columns of Y, wave w rows
// [16w, 16w+16) x 64 columns (4 MFMA tiles).  K (= ka) is walked 4 at a time.
__global__ __launch_bounds__(256) void gemm_nn_tall_kernel(const float* __restrict__ A,
                                                           int64_t lda, int64_t ka,
                                                           const double* __restrict__ B,
                                                           int64_t kb, int64_t n,
                                                           float* __restrict__ Y, int64_t ldy) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * 64 + wave * 16;
  const int64_t col0 = (int64_t)blockIdx.y * 64;
  const int64_t arow = row0 + l16;
  const bool rok = arow < n;
  const float* pa = A + (rok ? arow : 0) * lda;
  f64x4_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = (f64x4_t){0, 0, 0, 0};
  for (int64_t k0 = 0; k0 < ka; k0 += 16) {
    double av[4], bv[4][4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int64_t kk = k0 + 4 * s + kq;
      bool kok = kk < ka;
      av[s] = (rok && kok) ? (double)pa[kk] : 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int64_t col = col0 + 16 * j + l16;
        bv[s][j] = (kok && col < kb) ? B[kk * kb + col] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s][j], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int64_t r = row0 + kq + 4 * q, col = col0 + 16 * j + l16;  // D: row lane/16 + 4q
      if (r < n && col < kb) Y[r * ldy + col] = (float)acc[j][q];
    }
}

// ka % 16 == 0, kb % 64 == 0, A rows 16-byte aligned: 256 threads; block owns 256 rows x 64
// columns of Y, wave w rows [64w, 64w + 64): four 16-row blocks x four 16-column MFMA
// tiles, so each B element loaded feeds four MFMAs.  K is walked 16 at a time: lane group
// kq takes k0 + 4 kq .. + 3 (one 16-byte load of A per row block), MFMA s of the step
// k0 + 4 kq + s.  One wave per SIMD (the 16 accumulators take 128 registers), so the
// operands of step k0 + 16 are loaded into a second register set while step k0's MFMAs
// run (loop unrolled by two).  Loads are branch-free: rows past n read row n - 1 (their
// results are not stored), the prefetch past the last step reloads it.
constexpr int GT_RB = 4;
__global__ __launch_bounds__(256) void gemm_nn_tall_rb_kernel(const float* __restrict__ A,
                                                              int64_t lda, int64_t ka,
                                                              const double* __restrict__ B,
                                                              int64_t kb, int64_t n,
                                                              float* __restrict__ Y, int64_t ldy) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  // 1-D grid, XCD-aware: blocks go to the 8 XCDs round-robin by index, and an XCD's
  // consecutive blocks take the column tiles of one row block in turn, so the row block of
  // A (256 rows) is read from HBM once and from the XCD's L2 for the other column tiles
  const int64_t ncb = kb / 64, nrb = (n + 64 * GT_RB - 1) / (64 * GT_RB);
  const int64_t bid = blockIdx.x, j = bid / 8;
  const int64_t rblk = (j / ncb) * 8 + bid % 8;
  if (rblk >= nrb) return;
  const int64_t row0 = rblk * (64 * GT_RB) + wave * (16 * GT_RB);
  const int64_t col0 = (j % ncb) * 64;
  const float* pa[GT_RB];
#pragma unroll
  for (int rb = 0; rb < GT_RB; rb++) {
    const int64_t arow = row0 + 16 * rb + l16;
    pa[rb] = A + (arow < n ? arow : n - 1) * lda + 4 * kq;
  }
  const double* pb = B + (int64_t)(4 * kq) * kb + col0 + l16;
  f64x4_t acc[GT_RB][4];
#pragma unroll
  for (int rb = 0; rb < GT_RB; rb++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[rb][j] = (f64x4_t){0, 0, 0, 0};
  auto load = [&](int64_t k0, f32x4_t(&av)[GT_RB], double(&bv)[4][4]) {
#pragma unroll
    for (int sx = 0; sx < 4; sx++)
#pragma unroll
      for (int j = 0; j < 4; j++) bv[sx][j] = pb[(k0 + sx) * kb + 16 * j];
#pragma unroll
    for (int rb = 0; rb < GT_RB; rb++) av[rb] = *(const f32x4_t*)(pa[rb] + k0);
  };
  // the 16 accumulators in turn: an accumulator's next update is 16 MFMAs later
  auto step = [&](const f32x4_t(&av)[GT_RB], const double(&bv)[4][4]) {
#pragma unroll
    for (int sx = 0; sx < 4; sx++)
#pragma unroll
      for (int rb = 0; rb < GT_RB; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          acc[rb][j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[rb][sx], bv[sx][j],
                                                            acc[rb][j], 0, 0, 0);
  };
  f32x4_t aa[GT_RB], ab[GT_RB];
  double ba[4][4], bb[4][4];
  load(0, aa, ba);
  for (int64_t k0 = 0; k0 < ka; k0 += 32) {
    load(k0 + 16 < ka ? k0 + 16 : k0, ab, bb);
    step(aa, ba);
    if (k0 + 16 >= ka) break;
    load(k0 + 32 < ka ? k0 + 32 : k0 + 16, aa, ba);
    step(ab, bb);
  }
#pragma unroll
  for (int rb = 0; rb < GT_RB; rb++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        // D: row lane / 16 + 4 q of the 16-row block, column lane % 16 of tile j
        const int64_t r = row0 + 16 * rb + kq + 4 * q, col = col0 + 16 * j + l16;
        if (r < n) Y[r * ldy + col] = (float)acc[rb][j][q];
      }
}

tsvm_status gemm_nn_tall(Ctx* c, const float* A, int64_t lda, int64_t ka, const double* B,
                         int64_t kb, int64_t n, float* Y, int64_t ldy) {
  if (n <= 0 || kb <= 0) return TSVM_OK;
  if (ka % 16 == 0 && kb % 64 == 0 && lda % 4 == 0 && ((uintptr_t)A & 15) == 0) {
    // row blocks rounded up to whole rounds of the 8 XCDs (the blocks past n return at once)
    const int64_t nrb = ceil_div(n, 64 * GT_RB);
    const dim3 grid((unsigned)(round_up(nrb, 8) * (kb / 64)));
    hipLaunchKernelGGL(gemm_nn_tall_rb_kernel, grid, dim3(256), 0, c->stream, A, lda, ka, B, kb,
                       n, Y, ldy);
  } else {
    dim3 grid((unsigned)ceil_div(n, 64), (unsigned)ceil_div(kb, 64));
    hipLaunchKernelGGL(gemm_nn_tall_kernel, grid, dim3(256), 0, c->stream, A, lda, ka, B, kb, n,
                       Y, ldy);
  }
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

}  // namespace tsvm
