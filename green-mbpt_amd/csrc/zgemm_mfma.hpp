// Complex FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C[b](m, n) (+)= alpha * sum_k opA(A[b](m, k)) * opB(B[b](k, n))
//
// The GW contractions address their operands through two-level strides: a K index that
// runs over (k-point, orbital pair) and an N index that runs over (k-point, orbital pair),
// so every operand element is located by
//   A(m, k) = A + b*sAb + m*sAm + (k / kin)*sAko + (k % kin)*sAki
//   B(k, n) = B + b*sBb + (k / kin)*sBko + (k % kin)*sBki + (n / nin)*sBno + (n % nin)*sBni
//   C(m, n) = C + b*sCb + m*sCm + (n / nin)*sCno + (n % nin)*sCni
// which lets one kernel run the P0 contraction, the Sigma contraction, the IR transforms
// and the Dyson products straight out of their resident layouts.
//
// Tiling: 256 threads = 4 waves of 64; block tile 64x64 complex, K tile 16; each wave owns a
// 32x32 complex sub-tile = 2x2 MFMA blocks x (re, im) accumulators.  Operands are staged
// through LDS as separate real / imaginary planes (k-major), global loads for the next K tile
// are issued before the MFMAs of the current one.  A complex product is four real MFMAs.
#pragma once

#include "common.hpp"

namespace gwhip {

struct ZGemmArgs {
  int M, N, K, batch;
  int kin, nin;
  const double2* A;
  int64_t sAb, sAm, sAko, sAki;
  const double2* B;
  int64_t sBb, sBko, sBki, sBno, sBni;
  double2* C;
  int64_t sCb, sCm, sCno, sCni;
  double2 alpha;
  int accumulate;  // 1: C += alpha*AB, 0: C = alpha*AB
  int conjA, conjB;
  int tiles_m, tiles_n;
  int64_t nblocks;      // batch * tiles_m * tiles_n
  int64_t per_xcd;      // ceil(nblocks / 8)
};

constexpr int ZG_BM = 64, ZG_BN = 64, ZG_BK = 16, ZG_PAD = 2;

// A_KC: A is contiguous along k (thread loads run along k); otherwise along m.
// B_KC: B is contiguous along k; otherwise along n.
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256) void zgemm_f64_kernel(ZGemmArgs p) {
  __shared__ double As_r[ZG_BK][ZG_BM + ZG_PAD], As_i[ZG_BK][ZG_BM + ZG_PAD];
  __shared__ double Bs_r[ZG_BK][ZG_BN + ZG_PAD], Bs_i[ZG_BK][ZG_BN + ZG_PAD];

  // XCD-aware tile order: consecutive hardware workgroup ids go to different XCDs, so give
  // each XCD a contiguous range of logical tiles (tiles of one batch entry share operands).
  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return;
  const int tiles = p.tiles_m * p.tiles_n;
  const int64_t b = logical / tiles;
  const int rem = (int)(logical - b * tiles);
  const int tm = rem / p.tiles_n, tn = rem - (rem / p.tiles_n) * p.tiles_n;
  const int m0 = tm * ZG_BM, n0 = tn * ZG_BN;

  const double2* __restrict__ A = p.A + b * p.sAb;
  const double2* __restrict__ B = p.B + b * p.sBb;
  double2* __restrict__ C = p.C + b * p.sCb;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const double sa = p.conjA ? -1.0 : 1.0, sb = p.conjB ? -1.0 : 1.0;

  // load coordinates (4 elements per thread per operand)
  int a_m[4], a_k[4], b_k[4], b_n[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (A_KC) { a_m[j] = tid >> 2; a_k[j] = (tid & 3) * 4 + j; }
    else      { a_k[j] = tid >> 4; a_m[j] = (tid & 15) * 4 + j; }
    if (B_KC) { b_n[j] = tid >> 2; b_k[j] = (tid & 3) * 4 + j; }
    else      { b_k[j] = tid >> 4; b_n[j] = (tid & 15) * 4 + j; }
  }
  // per-thread column offsets of B (n is fixed per thread across K tiles)
  int64_t b_noff[4];
  bool b_nok[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + b_n[j];
    b_nok[j] = n < p.N;
    const int no = n / p.nin;
    b_noff[j] = (int64_t)no * p.sBno + (int64_t)(n - no * p.nin) * p.sBni;
  }

  double2 ra[4], rb[4];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + a_m[j], k = k0 + a_k[j];
      double2 v = make_double2(0.0, 0.0);
      if (m < p.M && k < p.K) {
        const int ko = k / p.kin;
        v = A[(int64_t)m * p.sAm + (int64_t)ko * p.sAko + (int64_t)(k - ko * p.kin) * p.sAki];
      }
      ra[j] = v;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + b_k[j];
      double2 v = make_double2(0.0, 0.0);
      if (b_nok[j] && k < p.K) {
        const int ko = k / p.kin;
        v = B[(int64_t)ko * p.sBko + (int64_t)(k - ko * p.kin) * p.sBki + b_noff[j]];
      }
      rb[j] = v;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      As_r[a_k[j]][a_m[j]] = ra[j].x;
      As_i[a_k[j]][a_m[j]] = sa * ra[j].y;
      Bs_r[b_k[j]][b_n[j]] = rb[j].x;
      Bs_i[b_k[j]][b_n[j]] = sb * rb[j].y;
    }
  };

  v4d acc_r[2][2], acc_i[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      acc_r[x][y] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_i[x][y] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lrow = lane & 15, lk = lane >> 4;

  load_tiles(0);
  for (int k0 = 0; k0 < p.K; k0 += ZG_BK) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (k0 + ZG_BK < p.K) load_tiles(k0 + ZG_BK);
#pragma unroll
    for (int ks = 0; ks < ZG_BK; ks += 4) {
      const int kr = ks + lk;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        ar[x] = As_r[kr][wm + x * 16 + lrow];
        ai[x] = As_i[kr][wm + x * 16 + lrow];
        br[x] = Bs_r[kr][wn + x * 16 + lrow];
        bi[x] = Bs_i[kr][wn + x * 16 + lrow];
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          acc_r[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], acc_r[x][y], 0, 0, 0);
          acc_r[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[x], bi[y], acc_r[x][y], 0, 0, 0);
          acc_i[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], acc_i[x][y], 0, 0, 0);
          acc_i[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], acc_i[x][y], 0, 0, 0);
        }
    }
  }

  // epilogue: register r of lane l holds row (l/16) + 4*r, column l%16 of each 16x16 block
  // (v_mfma_f64_16x16x4f64 result layout, checked by the identity-probe test)
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int n = n0 + wn + y * 16 + lrow;
      if (n >= p.N) continue;
      const int no = n / p.nin;
      const int64_t coff = (int64_t)no * p.sCno + (int64_t)(n - no * p.nin) * p.sCni;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + x * 16 + lk + 4 * r;
        if (m >= p.M) continue;
        const double cr = acc_r[x][y][r], ci = acc_i[x][y][r];
        double2 v = make_double2(p.alpha.x * cr - p.alpha.y * ci, p.alpha.x * ci + p.alpha.y * cr);
        double2* dst = C + (int64_t)m * p.sCm + coff;
        if (p.accumulate) {
          const double2 o = *dst;
          v.x += o.x;
          v.y += o.y;
        }
        *dst = v;
      }
    }
}

// Host-side launcher.  Fills tile counts and the XCD remap, checks sizes.
inline void zgemm_f64(ZGemmArgs p, bool a_kc, bool b_kc, hipStream_t stream) {
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return;
  if (p.kin <= 0 || p.nin <= 0) throw std::runtime_error("zgemm_f64: kin/nin must be positive");
  p.tiles_m = (int)ceil_div(p.M, ZG_BM);
  p.tiles_n = (int)ceil_div(p.N, ZG_BN);
  p.nblocks = (int64_t)p.batch * p.tiles_m * p.tiles_n;
  p.per_xcd = ceil_div(p.nblocks, 8);
  const int64_t grid = p.per_xcd * 8;
  if (grid > 0x7fffffffLL) throw std::runtime_error("zgemm_f64: grid too large");
  dim3 g((unsigned)grid), blk(256);
  if (a_kc && b_kc)
    hipLaunchKernelGGL((zgemm_f64_kernel<true, true>), g, blk, 0, stream, p);
  else if (a_kc && !b_kc)
    hipLaunchKernelGGL((zgemm_f64_kernel<true, false>), g, blk, 0, stream, p);
  else if (!a_kc && b_kc)
    hipLaunchKernelGGL((zgemm_f64_kernel<false, true>), g, blk, 0, stream, p);
  else
    hipLaunchKernelGGL((zgemm_f64_kernel<false, false>), g, blk, 0, stream, p);
  GWHIP_CHECK(hipGetLastError());
}

}  // namespace gwhip
