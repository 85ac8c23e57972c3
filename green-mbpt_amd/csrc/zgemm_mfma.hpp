// Complex FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C[b](m, n) (+)= alpha * sum_k opA(A[b](m, k)) * opB(B[b](k, n))
//
// The GW contractions address their operands through two-level strides: a K index that
// runs over (k-point, orbital pair) and an N index that runs over (k-point, orbital pair),
// so every operand element is located by
//   A(m, k) = A + b*sAb + m*sAm + (k / kin)*sAko + (k % kin)*sAki
//   B(k, n) = B + b*sBb + (k / kin)*sBko + (k % kin)*sBki + (n / nin)*sBno + (n % nin)*sBni
// (with nin2 > 0 the inner part splits once more: ni = n % nin -> (ni / nin2)*sBni + (ni % nin2)*sBni2,
// which reads a per-block transposed operand in place)
//   C(m, n) = C + b*sCb + m*sCm + (n / nin)*sCno + (n % nin)*sCni
// which lets one kernel run the P0 contraction, the Sigma contraction, the IR transforms
// and the Dyson products straight out of their resident layouts.
//
// Tiling: 256 threads = 4 waves in a 2x2 grid; block tile BM x BN complex, K tile 16; each
// wave owns a (BM/2) x (BN/2) sub-tile of 16x16 MFMA blocks with separate real/imaginary
// accumulators.  Operands are staged through LDS as real/imaginary planes (k-major); the
// global loads of the next K tile are in flight during the MFMAs of the current one.  Each
// thread walks its operand addresses incrementally along K, so the hot loop carries no
// integer division.  A complex product is four real MFMAs.
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
  int nin2;          // 0: B's inner N index is one level; > 0: split as above
  int64_t sBni2;
  double2* C;
  int64_t sCb, sCm, sCno, sCni;
  double2 alpha;
  int accumulate;  // 1: C += alpha*AB, 0: C = alpha*AB
  int conjA, conjB;
  int upper;       // 1: skip block tiles that lie wholly below the diagonal (Hermitian products)
  int tiles_m, tiles_n;
  int64_t nblocks;   // batch * tiles_m * tiles_n * ksplit
  int64_t per_xcd;   // ceil(nblocks / 8)
  int ksplit;        // K cut into ksplit slices of kslice (a multiple of the K step), one block each;
  int kslice;        // slice s writes its partial product at C + s * sCk (0 or 1: no split)
  int64_t sCk;
};

constexpr int ZG_BK = 16, ZG_PAD = 2;

// Running position (ko, ki) of one operand element along K with its address offset.
struct KWalk {
  int ki;
  int64_t off;
  __device__ __forceinline__ void init(int k, int kin, int64_t sko, int64_t ski) {
    const int ko = k / kin;
    ki = k - ko * kin;
    off = (int64_t)ko * sko + (int64_t)ki * ski;
  }
  __device__ __forceinline__ void advance(int kin, int64_t sko, int64_t ski) {
    ki += ZG_BK;
    off += (int64_t)ZG_BK * ski;
    while (ki >= kin) {
      ki -= kin;
      off += sko - (int64_t)kin * ski;
    }
  }
};

template <int BM, int BN, bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, (BM * BN > 4096) ? 1 : 2) void zgemm_f64_kernel(ZGemmArgs p) {
  constexpr int BK = ZG_BK;
  constexpr int WM = BM / 2, WN = BN / 2;        // per-wave tile
  constexpr int MB = WM / 16, NB = WN / 16;      // 16x16 blocks per wave
  constexpr int EA = BM * BK / 256, EB = BN * BK / 256;  // elements per thread
  __shared__ double As_r[BK][BM + ZG_PAD], As_i[BK][BM + ZG_PAD];
  __shared__ double Bs_r[BK][BN + ZG_PAD], Bs_i[BK][BN + ZG_PAD];

  // XCD-aware tile order: consecutive hardware workgroup ids go to different XCDs, so give
  // each XCD a contiguous range of logical tiles (tiles of one batch entry share operands).
  const int64_t hw = blockIdx.x;
  int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return;
  int kbeg = 0, kend = p.K;
  int64_t cks = 0;
  if (p.ksplit > 1) {  // K slice of this block
    const int64_t ks = logical % p.ksplit;
    logical /= p.ksplit;
    kbeg = (int)ks * p.kslice;
    kend = kbeg + p.kslice < p.K ? kbeg + p.kslice : p.K;
    cks = ks * p.sCk;
  }
  const int tiles = p.tiles_m * p.tiles_n;
  const int64_t b = logical / tiles;
  const int rem = (int)(logical - b * tiles);
  const int tm = rem / p.tiles_n, tn = rem - (rem / p.tiles_n) * p.tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;
  if (p.upper && m0 >= n0 + BN) return;

  const double2* __restrict__ A = p.A + b * p.sAb;
  const double2* __restrict__ B = p.B + b * p.sBb;
  double2* __restrict__ C = p.C + b * p.sCb + cks;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const double sa = p.conjA ? -1.0 : 1.0, sb = p.conjB ? -1.0 : 1.0;

  // thread -> tile coordinates of its EA (EB) elements
  int a_m[EA], a_k[EA], b_k[EB], b_n[EB];
#pragma unroll
  for (int j = 0; j < EA; ++j) {
    if (A_KC) {  // runs of 4 consecutive k
      const int id = tid + (j / 4) * 256;
      a_m[j] = id / (BK / 4);
      a_k[j] = (id % (BK / 4)) * 4 + (j % 4);
    } else {     // runs of 4 consecutive m
      const int id = tid + (j / 4) * 256;
      a_k[j] = id / (BM / 4);
      a_m[j] = (id % (BM / 4)) * 4 + (j % 4);
    }
  }
#pragma unroll
  for (int j = 0; j < EB; ++j) {
    if (B_KC) {
      const int id = tid + (j / 4) * 256;
      b_n[j] = id / (BK / 4);
      b_k[j] = (id % (BK / 4)) * 4 + (j % 4);
    } else {
      const int id = tid + (j / 4) * 256;
      b_k[j] = id / (BN / 4);
      b_n[j] = (id % (BN / 4)) * 4 + (j % 4);
    }
  }
  // fixed row / column parts of the addresses
  int64_t a_moff[EA], b_noff[EB];
  bool a_ok[EA], b_ok[EB];
  KWalk ka[EA], kb[EB];
#pragma unroll
  for (int j = 0; j < EA; ++j) {
    const int m = m0 + a_m[j];
    a_ok[j] = m < p.M;
    a_moff[j] = (int64_t)(a_ok[j] ? m : 0) * p.sAm;
    ka[j].init(kbeg + a_k[j], p.kin, p.sAko, p.sAki);
  }
#pragma unroll
  for (int j = 0; j < EB; ++j) {
    const int n = n0 + b_n[j];
    b_ok[j] = n < p.N;
    const int nn = b_ok[j] ? n : 0;
    const int no = nn / p.nin, ni = nn - no * p.nin;
    if (p.nin2 > 0) {
      const int nh = ni / p.nin2;
      b_noff[j] = (int64_t)no * p.sBno + (int64_t)nh * p.sBni + (int64_t)(ni - nh * p.nin2) * p.sBni2;
    } else {
      b_noff[j] = (int64_t)no * p.sBno + (int64_t)ni * p.sBni;
    }
    kb[j].init(kbeg + b_k[j], p.kin, p.sBko, p.sBki);
  }

  double2 ra[EA], rb[EB];
  auto load_tiles = [&](int k0) {
    const bool full = k0 + BK <= kend;
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const bool ok = a_ok[j] && (full || k0 + a_k[j] < kend);
      ra[j] = ok ? A[a_moff[j] + ka[j].off] : make_double2(0.0, 0.0);
      ka[j].advance(p.kin, p.sAko, p.sAki);
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const bool ok = b_ok[j] && (full || k0 + b_k[j] < kend);
      rb[j] = ok ? B[kb[j].off + b_noff[j]] : make_double2(0.0, 0.0);
      kb[j].advance(p.kin, p.sBko, p.sBki);
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      As_r[a_k[j]][a_m[j]] = ra[j].x;
      As_i[a_k[j]][a_m[j]] = sa * ra[j].y;
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      Bs_r[b_k[j]][b_n[j]] = rb[j].x;
      Bs_i[b_k[j]][b_n[j]] = sb * rb[j].y;
    }
  };

  v4d acc_r[MB][NB], acc_i[MB][NB];
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      acc_r[x][y] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_i[x][y] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int lrow = lane & 15, lk = lane >> 4;

  load_tiles(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (k0 + BK < kend) load_tiles(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + lk;
      double ar[MB], ai[MB], br[NB], bi[NB];
#pragma unroll
      for (int x = 0; x < MB; ++x) {
        ar[x] = As_r[kr][wm + x * 16 + lrow];
        ai[x] = As_i[kr][wm + x * 16 + lrow];
      }
#pragma unroll
      for (int y = 0; y < NB; ++y) {
        br[y] = Bs_r[kr][wn + y * 16 + lrow];
        bi[y] = Bs_i[kr][wn + y * 16 + lrow];
      }
      // independent products first, the dependent second halves after
#pragma unroll
      for (int x = 0; x < MB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          acc_r[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], acc_r[x][y], 0, 0, 0);
          acc_i[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], acc_i[x][y], 0, 0, 0);
        }
#pragma unroll
      for (int x = 0; x < MB; ++x) {
        const double nai = -ai[x];
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          acc_r[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi[y], acc_r[x][y], 0, 0, 0);
          acc_i[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], acc_i[x][y], 0, 0, 0);
        }
      }
    }
  }

  // epilogue: register r of lane l holds row (l/16) + 4*r, column l%16 of each 16x16 block
  // (v_mfma_f64_16x16x4f64 result layout, checked by the identity-probe test)
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
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

// Tile shapes: 0 = 64x64, 1 = 128x64, 2 = 64x128.
template <int BM, int BN>
inline void zgemm_launch_shape(const ZGemmArgs& p, bool a_kc, bool b_kc, dim3 g, hipStream_t stream) {
  dim3 blk(256);
  if (a_kc && b_kc)
    hipLaunchKernelGGL((zgemm_f64_kernel<BM, BN, true, true>), g, blk, 0, stream, p);
  else if (a_kc && !b_kc)
    hipLaunchKernelGGL((zgemm_f64_kernel<BM, BN, true, false>), g, blk, 0, stream, p);
  else if (!a_kc && b_kc)
    hipLaunchKernelGGL((zgemm_f64_kernel<BM, BN, false, true>), g, blk, 0, stream, p);
  else
    hipLaunchKernelGGL((zgemm_f64_kernel<BM, BN, false, false>), g, blk, 0, stream, p);
}

// Host-side launcher.  Fills tile counts and the XCD remap, checks sizes.
inline void zgemm_f64(ZGemmArgs p, bool a_kc, bool b_kc, hipStream_t stream, int shape = 0) {
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return;
  if (p.kin <= 0 || p.nin <= 0) throw std::runtime_error("zgemm_f64: kin/nin must be positive");
  if (p.ksplit > 1 && (p.kslice <= 0 || p.kslice % ZG_BK != 0 || (int64_t)p.ksplit * p.kslice < p.K))
    throw std::runtime_error("zgemm_f64: bad K split");
  const int BM = shape == 1 ? 128 : 64, BN = shape == 2 ? 128 : 64;
  p.tiles_m = (int)ceil_div(p.M, BM);
  p.tiles_n = (int)ceil_div(p.N, BN);
  p.nblocks = (int64_t)p.batch * p.tiles_m * p.tiles_n * (p.ksplit > 1 ? p.ksplit : 1);
  p.per_xcd = ceil_div(p.nblocks, 8);
  const int64_t grid = p.per_xcd * 8;
  if (grid > 0x7fffffffLL) throw std::runtime_error("zgemm_f64: grid too large");
  const dim3 g((unsigned)grid);
  if (shape == 1)
    zgemm_launch_shape<128, 64>(p, a_kc, b_kc, g, stream);
  else if (shape == 2)
    zgemm_launch_shape<64, 128>(p, a_kc, b_kc, g, stream);
  else
    zgemm_launch_shape<64, 64>(p, a_kc, b_kc, g, stream);
  GWHIP_CHECK(hipGetLastError());
}

// Sum of the ksplit partial products Part[s][b][m][n] (n fastest) into
// C(b, m, n) = alpha * sum + (accumulate ? C : 0), with C addressed as in ZGemmArgs.
__global__ void zgemm_splitk_reduce_kernel(ZGemmArgs p, const double2* __restrict__ part) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t mn = (int64_t)p.M * p.N;
  if (e >= mn * p.batch) return;
  const int64_t b = e / mn;
  const int m = (int)((e - b * mn) / p.N), n = (int)(e - b * mn - (int64_t)m * p.N);
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < p.ksplit; ++s) {
    const double2 v = part[s * p.sCk + e];
    sr += v.x;
    si += v.y;
  }
  const int no = n / p.nin;
  double2* dst = p.C + b * p.sCb + (int64_t)m * p.sCm + (int64_t)no * p.sCno + (int64_t)(n - no * p.nin) * p.sCni;
  double2 v = make_double2(p.alpha.x * sr - p.alpha.y * si, p.alpha.x * si + p.alpha.y * sr);
  if (p.accumulate) {
    const double2 o = *dst;
    v.x += o.x;
    v.y += o.y;
  }
  *dst = v;
}

// C (+)= alpha A B with K cut into `ksplit` slices computed by separate blocks (for few, long-K
// products); part holds ksplit * batch * M * N elements.  The slices are summed in slice order,
// so the result does not depend on the block schedule.
inline void zgemm_f64_splitk(ZGemmArgs p, int ksplit, double2* part, bool a_kc, bool b_kc, hipStream_t stream) {
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return;
  if (ksplit < 2) {
    zgemm_f64(p, a_kc, b_kc, stream);
    return;
  }
  ZGemmArgs q = p;
  q.kslice = (int)(ceil_div(ceil_div(p.K, ksplit), ZG_BK) * ZG_BK);
  q.ksplit = (int)ceil_div(p.K, q.kslice);
  q.sCk = (int64_t)p.batch * p.M * p.N;
  q.C = part;
  q.sCb = (int64_t)p.M * p.N;
  q.sCm = p.N;
  q.sCno = p.nin;
  q.sCni = 1;
  q.alpha = make_double2(1.0, 0.0);
  q.accumulate = 0;
  q.upper = 0;
  zgemm_f64(q, a_kc, b_kc, stream);
  ZGemmArgs r = p;
  r.ksplit = q.ksplit;
  r.sCk = q.sCk;
  const int64_t n = (int64_t)p.batch * p.M * p.N;
  hipLaunchKernelGGL(zgemm_splitk_reduce_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, r, part);
  GWHIP_CHECK(hipGetLastError());
}

}  // namespace gwhip
