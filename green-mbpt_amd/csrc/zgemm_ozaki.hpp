// Split-precision complex GEMM: FP64 operands cut into int8 digit slices, products on the int8
// matrix cores (v_mfma_i32_16x16x64_i8) with exact int32 accumulation inside a K chunk and FP64
// accumulation across chunks.  This is the fixed-point slicing of Ozaki et al. (Numer.
// Algorithms 59, 2012; Ootomo, Ozaki & Yokota, IJHPCA 2024 for integer matrix units), with
// balanced digits:
//
//   x = 2^e / 127 * sum_{s<NS} x_s 254^{-s} + r,   x_s in [-127, 127],  |r| <= 2^e / 127 * 254^{-NS} / 2
//
// with one power-of-two exponent e per (row, K chunk) of A and per (column, K chunk) of B
// (|Re x|, |Im x| < 2^e over the chunk).  Digits come from round-to-nearest, so each carries
// log2(254) ~ 7.99 bits (truncated radix-128 digits carry 7): four slices keep 32 bits, five 40.
// Every slice product is an integer; pairs with s + t = d share the weight 254^{-d} and one int32
// accumulator per real product plane, exact while 2 * NS * kc * 127^2 < 2^31 (split_kc_limit);
// pairs with s + t >= NS are dropped.
//
// Slice store (int8), for each row (A) or column (B):
//   [chunk][plane][slice][kc]   planes Re, Im
// (B is stored already conjugated when the contraction needs conj(B)), so
//   Re C = Re A Re B + Im A (-Im B),   Im C = Re A Im B + Im A Re B,
// with -Im B (32 x 64 tile) or -Im A (64 x 64 tile) negated in registers (digits lie in [-127, 127],
// so negation is exact).  Karatsuba's
// three-product form needs a third (Re + Im) plane and a third accumulator set: at five slices the
// 32 x 64 tile then runs out of registers (256 VGPRs and a 117-register spill per lane at two
// waves per SIMD), so the four-product form is kept.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace gwhip {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int OZ_PLANES = 2;
constexpr double OZ_FIRST = 127.0, OZ_RADIX = 254.0;

// 254^{-d}, d < 8, correctly rounded
__device__ __forceinline__ double oz_weight(int d) {
  constexpr double w[8] = {1.0, 1.0 / 254.0, 1.0 / 64516.0, 1.0 / 16387064.0, 1.0 / 4162314256.0,
                           1.0 / 1057227821024.0, 1.0 / 268535866540096.0, 1.0 / 68208110101184384.0};
  return w[d];
}

// Balanced digits of v with |v| <= 127 (v = x 127 / 2^e): d_s = rint(v), v <- (v - d_s) 254.  Every
// |v| stays <= 127, so every digit fits in [-127, 127].
template <int NS>
__device__ __forceinline__ void oz_digits(double v, int (&d)[NS]) {
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double q = rint(v);
    d[s] = (int)q;
    v = (v - q) * OZ_RADIX;
  }
}

// ------------------------------------------------------------------------------ slicing
// Row r = r_hi * row_div + r_lo of a strided FP64 complex source; element kk of chunk c at
//   src + r_hi*s_rhi + r_lo*s_rlo + c*s_c + kk*s_k,   kk < klen (zero up to kc);
// with row_div2 > 0 the low part splits once more, r_lo = a * row_div2 + b at a*s_rlo + b*s_rlo2
// (rows that run over a per-block transposed source).
struct SliceArgs {
  const double2* src;
  int64_t nrows;
  int64_t row_div, s_rhi, s_rlo;
  int64_t row_div2, s_rlo2;
  int nchunk;
  int64_t s_c;
  int klen;
  int64_t s_k;
  int kc;        // padded chunk length (multiple of 64)
  int conj;      // slice conj(x)
  int8_t* dst;   // [row][chunk][plane][slice][kc]
  int* exps;     // [row][chunk]
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// The planes' digits of one complex element (already conjugated if asked) with exponent e: Re and Im
// scaled by sc = 127 / 2^e.
template <int NS>
__device__ __forceinline__ void oz_element_digits(double re, double im, double sc, int (&dr)[NS], int (&di)[NS]) {
  oz_digits<NS>(re * sc, dr);
  oz_digits<NS>(im * sc, di);
}

// One wave per (row, chunk); each lane slices 4 consecutive elements per step.
template <int NS>
__global__ __launch_bounds__(256) void slice_rows_kernel(SliceArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t task = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (task >= a.nrows * a.nchunk) return;
  const int64_t row = task / a.nchunk;
  const int c = (int)(task - row * a.nchunk);
  const int64_t rhi = row / a.row_div, rlo = row - rhi * a.row_div;
  int64_t roff = rlo * a.s_rlo;
  if (a.row_div2 > 0) {
    const int64_t rh = rlo / a.row_div2;
    roff = rh * a.s_rlo + (rlo - rh * a.row_div2) * a.s_rlo2;
  }
  const double2* s = a.src + rhi * a.s_rhi + roff + (int64_t)c * a.s_c;
  double m = 0.0;
  for (int kk = lane; kk < a.klen; kk += 64) {
    const double2 v = s[(int64_t)kk * a.s_k];
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
  m = wave_max(m);
  int e = 0;
  if (m > 0.0) (void)frexp(m, &e);  // m < 2^e
  if (lane == 0) a.exps[task] = e;
  const double sc = ldexp(OZ_FIRST, -e);
  const double sgn = a.conj ? -1.0 : 1.0;
  int8_t* out = a.dst + task * (int64_t)OZ_PLANES * NS * a.kc;
  for (int base = 0; base < a.kc; base += 256) {
    const int k0 = base + 4 * lane;
    if (k0 >= a.kc) continue;
    int wr[NS], wi[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) { wr[t] = 0; wi[t] = 0; }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = k0 + j;
      double2 v = make_double2(0.0, 0.0);
      if (kk < a.klen) v = s[(int64_t)kk * a.s_k];
      int dr[NS], di[NS];
      oz_element_digits<NS>(v.x, sgn * v.y, sc, dr, di);
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        wr[t] |= (dr[t] & 0xff) << (8 * j);
        wi[t] |= (di[t] & 0xff) << (8 * j);
      }
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      *reinterpret_cast<int*>(out + (int64_t)(0 * NS + t) * a.kc + k0) = wr[t];
      *reinterpret_cast<int*>(out + (int64_t)(1 * NS + t) * a.kc + k0) = wi[t];
    }
  }
}

// ------------------------------------------------------------------------------ GEMM
// Largest K chunk (int8 elements per plane-slice, a multiple of 64) whose int32 accumulation stays
// exact with ns slices: an accumulator of slice level d sums at most ns slice pairs of two product
// planes, each |x y| <= 127^2 < 2^14, over kc elements, so 2 * ns * 2^14 * kc < 2^31.
inline int64_t split_kc_limit(int ns) { return ((((int64_t)1 << 31) - 1) / ((int64_t)2 * ns << 14)) / 64 * 64; }

// Byte-wise two's complement negation of 16 int8 values in [-127, 127]:
// -x = ~x + 1 per byte, the +1 kept inside each byte by adding it to the low seven bits.
__device__ __forceinline__ v4i neg_i8x16(v4i w) {
  v4i r;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const unsigned n = ~(unsigned)w[i];
    r[i] = (int)(((n & 0x7F7F7F7Fu) + 0x01010101u) ^ (n & 0x80808080u));
  }
  return r;
}

struct OzGemmArgs {
  int M, N, batch;
  int kc, nchunk;        // K = nchunk * kc (int8 elements per plane-slice)
  int nin;               // N split for B / C columns
  // A(row m, chunk c): slices at A + b*sAb + m*sAm + c*sAc, exponent eA[b*seAb + m*seAm + c]
  const int8_t* A;
  int64_t sAb, sAm, sAc;
  const int* eA;
  int64_t seAb, seAm;
  // B(col n, chunk c): slices at B + b*sBb + (n/nin)*sBno + (n%nin)*sBni + c*sBc,
  //                    exponent eB[b*seBb + (n/nin)*seBno + (n%nin)*seBni + c]
  const int8_t* B;
  int64_t sBb, sBno, sBni, sBc;
  const int* eB;
  int64_t seBb, seBno, seBni;
  double2* C;
  int64_t sCb, sCm, sCno, sCni;
  double2 alpha;
  int accumulate;
  int tiles_m, tiles_n;
  int64_t nblocks, per_xcd;
  int group_m;  // row tiles (over batch x M) swept together across all column tiles; 0 = heuristic
  int upper;    // 1: skip block tiles that lie wholly below the diagonal (Hermitian products)
  int tile;     // block tile: 2 = 32 x 64 of 16 x 16 x 64 MFMAs (zgemm_ozaki_pipe_kernel), 4 = 64 x 64 of
                // 32 x 32 x 32 MFMAs (zgemm_ozaki32_kernel); 0 = 2
};

// Output layout of v_mfma_i32_16x16x64_i8: register r of lane l holds row
// OZ_ROW(l/16, r), column l%16 (checked by the identity-probe test).
#define OZ_ROW(g, r) (4 * (g) + (r))

// Block tile 32 x 64 (complex); the four waves split N, each owning 32 x 16 = two 16x16 blocks.
// The products take the four-product form with -Im B: Re C = Re A Re B + Im A (-Im B), negated once
// per K step for the NS B fragments (negating -Im A instead costs one negation per slice and row
// block).  ONE (a single K chunk, as the Sigma contraction has): the int32 -> FP64 conversion runs once
// after the loop, so no FP64 sums are live inside it (156 instead of 220 VGPRs at four slices: three
// waves per SIMD instead of two).
// A (32 rows, shared by all waves) is staged through a double-buffered LDS panel one K step (64
// bytes per plane-slice row) at a time, its four 16-byte pieces XOR-swizzled by (r >> 2) & 3 so
// that the 16 rows of one fragment read fall on distinct banks; each LDS store covers 16 whole rows
// of one plane-slice.  Each wave reads its own 16 B columns straight from global memory into MFMA
// fragments.  Software pipeline: while the MFMAs of K step g run, the A pieces and B fragments of
// step g + 1 are in flight into registers; A goes to the other half of the LDS panel after the
// multiplies (one barrier per K step); fb[plane][t] is last used by slice s = NS - 1 - t, after
// which its registers take the next step's fragment.
constexpr int OZ_BM = 32, OZ_BN = 64, OZ_KS = 64;

__device__ __forceinline__ int oz_lds_off(int ps, int rows, int r, int piece) {
  return (ps * rows + r) * 64 + 16 * (piece ^ ((r >> 2) & 3));
}

template <int NS, bool ONE>
__global__ __launch_bounds__(256, 2) void zgemm_ozaki_pipe_kernel(OzGemmArgs p) {
  constexpr int NPS = OZ_PLANES * NS;          // plane-slices per row
  constexpr int NPC = NPS / 2;                 // staged per thread and K step (two parities)
  constexpr int BUF = NPS * OZ_BM * 64;
  extern __shared__ __attribute__((aligned(16))) int8_t oz_lds[];

  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return;
  // Each XCD takes a contiguous range of logical tiles; within it, groups of group_m row tiles
  // (rows run over batch x M) sweep every column tile, row fastest, so a group's A panels stay in
  // that XCD's L2 while the B panels stream through it once per group.
  const int64_t rows = (int64_t)p.batch * p.tiles_m;
  const int64_t in_group = (int64_t)p.group_m * p.tiles_n;
  const int64_t grp = logical / in_group;
  const int64_t first = grp * p.group_m;
  const int64_t gsize = rows - first < p.group_m ? rows - first : p.group_m;
  const int64_t idx = logical - grp * in_group;
  const int64_t rt = first + idx % gsize;
  const int tn = (int)(idx / gsize);
  const int64_t b = rt / p.tiles_m;
  const int tm = (int)(rt - b * p.tiles_m);
  const int m0 = tm * OZ_BM, n0 = tn * OZ_BN;
  if (p.upper && m0 >= n0 + OZ_BN) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lg = lane >> 4;
  const int wn = wave * 16;

  // A staging: thread -> (plane-slice parity, row, 16-byte piece), piece fastest.  Rows past M and
  // columns past N read the last valid row / column instead: they only feed results that are never
  // stored, and unconditional loads keep the memory-counter waits of the pipeline exact (a load
  // under a per-lane branch makes the compiler wait for every load in flight).
  const int a_row = (tid >> 2) & 31, a_piece = tid & 3, a_par = (tid >> 7) & 1;
  const int8_t* a_ptr = p.A + b * p.sAb + (int64_t)min(m0 + a_row, p.M - 1) * p.sAm + 16 * a_piece;
  const int a_lds = oz_lds_off(0, OZ_BM, a_row, a_piece);
  const int bcol = min(n0 + wn + lrow, p.N - 1);
  const int8_t* b_ptr;
  {
    const int no = bcol / p.nin;
    b_ptr = p.B + b * p.sBb + (int64_t)no * p.sBno + (int64_t)(bcol - no * p.nin) * p.sBni + 16 * lg;
  }
  const int nks = p.kc / OZ_KS, nsteps = p.nchunk * nks;
  auto b_at = [&](int g) -> const int8_t* {  // this lane's B bytes of K step g
    const int c = g / nks, ks = (g - c * nks) * OZ_KS;
    return b_ptr + (int64_t)c * p.sBc + ks;
  };

  // NPS = 2 NS is even, so the two parities cover every plane-slice: ps = a_par + 2 j < NPS
  static_assert(NPS % 2 == 0, "plane-slices come in pairs");
  v4i a_nx[NPC];
  auto load_a = [&](int g) {
    const int c = g / nks, ks = (g - c * nks) * OZ_KS;
#pragma unroll
    for (int j = 0; j < NPC; ++j) {
      const int ps = a_par + 2 * j;
      a_nx[j] = *reinterpret_cast<const v4i*>(a_ptr + (int64_t)c * p.sAc + (int64_t)ps * p.kc + ks);
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int j = 0; j < NPC; ++j) {
      const int ps = a_par + 2 * j;
      *reinterpret_cast<v4i*>(oz_lds + buf * BUF + a_lds + ps * OZ_BM * 64) = a_nx[j];
    }
  };

  double cr[2][4], ci[2][4];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int r = 0; r < 4; ++r) { cr[x][r] = 0.0; ci[x][r] = 0.0; }
  v4i ar[NS][2], ai[NS][2];  // [d][x] int32 accumulators (Re, Im)
#pragma unroll
  for (int d = 0; d < NS; ++d)
#pragma unroll
    for (int x = 0; x < 2; ++x) ar[d][x] = ai[d][x] = v4i{0, 0, 0, 0};

  // int32 -> FP64 with the (row, column, chunk) scale 2^(ea + eb) / 127^2; rows / columns past the
  // edge take the clamped ones' exponents (their results are not stored)
  auto convert = [&](int c) {
    const int no = bcol / p.nin;
    const int eb = p.eB[b * p.seBb + (int64_t)no * p.seBno + (int64_t)(bcol - no * p.nin) * p.seBni + c];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = min(m0 + 16 * x + OZ_ROW(lg, r), p.M - 1);
        const int ea = p.eA[b * p.seAb + (int64_t)m * p.seAm + c];
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int d = 0; d < NS; ++d) {  // exact integers times 254^-d
          sr = fma((double)ar[d][x][r], oz_weight(d), sr);
          si = fma((double)ai[d][x][r], oz_weight(d), si);
        }
        const double sc = ldexp(1.0 / (OZ_FIRST * OZ_FIRST), ea + eb);
        cr[x][r] = fma(sr, sc, cr[x][r]);
        ci[x][r] = fma(si, sc, ci[x][r]);
      }
  };

  v4i fb[OZ_PLANES][NS];
  load_a(0);
  store_a(0);
  {
    const int8_t* bc = b_at(0);
#pragma unroll
    for (int pl = 0; pl < OZ_PLANES; ++pl)
#pragma unroll
      for (int t = 0; t < NS; ++t) fb[pl][t] = *reinterpret_cast<const v4i*>(bc + (int64_t)(pl * NS + t) * p.kc);
  }
  __syncthreads();

  for (int g = 0; g < nsteps; ++g) {
    // the loads of the next step are issued unconditionally (the last step reloads its own data,
    // which nothing reads): a load under a branch makes the compiler wait for every load in flight
    const int gn = g + 1 < nsteps ? g + 1 : g;
    const int c = g / nks;
    const bool chunk_end = g - c * nks == nks - 1;
    load_a(gn);
    const int8_t* bn = b_at(gn);
    const int8_t* As = oz_lds + (g & 1) * BUF;
    v4i fbn[NS];  // -Im B of this K step
#pragma unroll
    for (int t = 0; t < NS; ++t) fbn[t] = neg_i8x16(fb[1][t]);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      v4i fa_r[2], fa_i[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int row = 16 * x + lrow;
        fa_r[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(0 * NS + s, OZ_BM, row, lg));
        fa_i[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(1 * NS + s, OZ_BM, row, lg));
      }
#pragma unroll
      for (int t = NS - 1 - s; t >= 0; --t) {
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          ar[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb[0][t], ar[s + t][x], 0, 0, 0);
          ai[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb[1][t], ai[s + t][x], 0, 0, 0);
          ar[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_i[x], fbn[t], ar[s + t][x], 0, 0, 0);
          ai[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_i[x], fb[0][t], ai[s + t][x], 0, 0, 0);
        }
      }
      {
        const int t = NS - 1 - s;
#pragma unroll
        for (int pl = 0; pl < OZ_PLANES; ++pl) fb[pl][t] = *reinterpret_cast<const v4i*>(bn + (int64_t)(pl * NS + t) * p.kc);
      }
    }
    if (!ONE && chunk_end) {
      convert(c);
#pragma unroll
      for (int d = 0; d < NS; ++d)
#pragma unroll
        for (int x = 0; x < 2; ++x) ar[d][x] = ai[d][x] = v4i{0, 0, 0, 0};
    }
    store_a((g + 1) & 1);  // on the last step into the buffer no wave reads any more
    __syncthreads();
  }
  if (ONE) convert(0);

  const int n = n0 + wn + lrow;
  if (n >= p.N) return;
  const int no = n / p.nin;
  const int64_t coff = (int64_t)no * p.sCno + (int64_t)(n - no * p.nin) * p.sCni;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + 16 * x + OZ_ROW(lg, r);
      if (m >= p.M) continue;
      const double vr = cr[x][r], vi = ci[x][r];
      double2 v = make_double2(p.alpha.x * vr - p.alpha.y * vi, p.alpha.x * vi + p.alpha.y * vr);
      double2* dst = p.C + b * p.sCb + (int64_t)m * p.sCm + coff;
      if (p.accumulate) {
        const double2 o = *dst;
        v.x += o.x;
        v.y += o.y;
      }
      *dst = v;
    }
}

// Logical tile of hardware block blockIdx.x in the XCD-aware, L2-grouped order: each XCD takes a
// contiguous range of logical tiles; within it, groups of group_m row tiles (rows run over
// batch x M) sweep every column tile, row fastest.  False for the padding blocks of the grid.
__device__ __forceinline__ bool oz_tile_of(const OzGemmArgs& p, int64_t& b, int& tm, int& tn) {
  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return false;
  const int64_t rows = (int64_t)p.batch * p.tiles_m;
  const int64_t in_group = (int64_t)p.group_m * p.tiles_n;
  const int64_t grp = logical / in_group;
  const int64_t first = grp * p.group_m;
  const int64_t gsize = rows - first < p.group_m ? rows - first : p.group_m;
  const int64_t idx = logical - grp * in_group;
  const int64_t rt = first + idx % gsize;
  tn = (int)(idx / gsize);
  b = rt / p.tiles_m;
  tm = (int)(rt - b * p.tiles_m);
  return true;
}

// ------------------------------------------------------------------------------ 32 x 32 x 32 tile
// Output layout of v_mfma_i32_32x32x32_i8: register r of lane l holds row OZ32_ROW(l/32, r),
// column l%32 (checked by the identity-probe test); each operand lane l holds row (A) / column (B)
// l%32 and K bytes 16 (l/32) .. +16 of the 32-byte K step.
#define OZ32_ROW(g, r) (8 * ((r) >> 2) + 4 * (g) + ((r)&3))
typedef int v16i __attribute__((ext_vector_type(16)));

// Block tile 64 x 64 (complex), four waves in a 2 x 2 grid of 32 x 32 wave tiles on the
// v_mfma_i32_32x32x32_i8 matrix cores, which keep their pipeline full with a handful of independent
// accumulators (the 16 x 16 x 64 form needs about sixteen in flight per wave).  Complex products in
// the four-product form on the two planes of the slice store:
//   Re C = Re A Re B + (-Im A) Im B,   Im C = Re A Im B + Im A Re B,
// -Im A negated bytewise in registers; per slice level two int32 accumulators of 16 registers.
// Operands go to LDS by direct global -> LDS loads (global_load_lds_dwordx4), double-buffered in two
// static arrays (distinct objects, so a wave reads one while the loads into the other are in
// flight); the loads of K step g + 1 leave at the start of step g, and the barrier closing step g
// waits for them.  A stage is a list of 1 KB chunks -- A's (row block of 32, plane-slice), then B's
// (column block of 32, plane-slice) -- each in MFMA fragment order, so one wave-wide load fills a
// chunk and a fragment read is one contiguous ds_read_b128 per lane.  Waves 0-1 load A's row blocks,
// waves 2-3 B's column blocks.  Rows past M / columns past N load the last valid row / column: they
// only feed results that are never stored.  Two blocks per CU.
constexpr int OZ32_BM = 64, OZ32_BN = 64, OZ32_KS = 32;

template <int NS>
__global__ __launch_bounds__(256, 2) void zgemm_ozaki32_kernel(OzGemmArgs p) {
  constexpr int NPU = 2 * NS;                   // plane-slices read: Re and Im
  constexpr int STAGE = 4 * NPU * 1024;         // 2 row blocks + 2 column blocks, NPU chunks each
  __shared__ __attribute__((aligned(1024))) int8_t oz_s0[STAGE];
  __shared__ __attribute__((aligned(1024))) int8_t oz_s1[STAGE];

  int64_t b;
  int tm, tn;
  if (!oz_tile_of(p, b, tm, tn)) return;
  const int m0 = tm * OZ32_BM, n0 = tn * OZ32_BN;
  if (p.upper && m0 >= n0 + OZ32_BN) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, lg = lane >> 5;
  const int wm = wave & 1, wn = wave >> 1;  // rows 32 wm .. +32, columns 32 wn .. +32

  // this lane's source row (waves 0-1: A row block `wave`) or column (waves 2-3: B column block)
  const bool loads_a = wave < 2;
  const int blk = wave & 1;
  const int8_t* src;
  int64_t s_chunk;
  if (loads_a) {
    const int m = min(m0 + 32 * blk + l32, p.M - 1);
    src = p.A + b * p.sAb + (int64_t)m * p.sAm + 16 * lg;
    s_chunk = p.sAc;
  } else {
    const int n = min(n0 + 32 * blk + l32, p.N - 1);
    const int no = n / p.nin;
    src = p.B + b * p.sBb + (int64_t)no * p.sBno + (int64_t)(n - no * p.nin) * p.sBni + 16 * lg;
    s_chunk = p.sBc;
  }
  const int chunk0 = (loads_a ? 0 : 2 * NPU) + blk * NPU;  // first of this wave's NPU chunks
  const int nks = p.kc / OZ32_KS, nsteps = p.nchunk * nks;
  auto load_stage = [&](int8_t* st, int g) {
    const int c = g / nks, ks = (g - c * nks) * OZ32_KS;
    const int8_t* s0 = src + (int64_t)c * s_chunk + ks;
#pragma unroll
    for (int ps = 0; ps < NPU; ++ps)
      __builtin_amdgcn_global_load_lds((const void*)(s0 + (int64_t)ps * p.kc),
                                       (__attribute__((address_space(3))) void*)(st + (chunk0 + ps) * 1024), 16, 0, 0);
  };

  double cr[16], ci[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) { cr[r] = 0.0; ci[r] = 0.0; }
  v16i ar[NS], ai[NS];  // [d] int32 accumulators (Re, Im) of slice level d
#pragma unroll
  for (int d = 0; d < NS; ++d)
#pragma unroll
    for (int r = 0; r < 16; ++r) { ar[d][r] = 0; ai[d][r] = 0; }

  const int bcol = min(n0 + 32 * wn + l32, p.N - 1);
  auto compute = [&](const int8_t* st, int g) {
    v4i fb[2][NS];
#pragma unroll
    for (int pl = 0; pl < 2; ++pl)
#pragma unroll
      for (int t = 0; t < NS; ++t)
        fb[pl][t] = *reinterpret_cast<const v4i*>(st + (2 * NPU + wn * NPU + pl * NS + t) * 1024 + 16 * lane);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const v4i fa_r = *reinterpret_cast<const v4i*>(st + (wm * NPU + 0 * NS + s) * 1024 + 16 * lane);
      const v4i fa_i = *reinterpret_cast<const v4i*>(st + (wm * NPU + 1 * NS + s) * 1024 + 16 * lane);
      const v4i fa_n = neg_i8x16(fa_i);
#pragma unroll
      for (int t = 0; t + s < NS; ++t) {
        ar[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa_r, fb[0][t], ar[s + t], 0, 0, 0);
        ai[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa_r, fb[1][t], ai[s + t], 0, 0, 0);
        ar[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa_n, fb[1][t], ar[s + t], 0, 0, 0);
        ai[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa_i, fb[0][t], ai[s + t], 0, 0, 0);
      }
    }
    const int c = g / nks;
    if (g - c * nks == nks - 1) {  // chunk end: int32 -> FP64 with the scale 2^(ea + eb) / 127^2
      const int no = bcol / p.nin;
      const int eb = p.eB[b * p.seBb + (int64_t)no * p.seBno + (int64_t)(bcol - no * p.nin) * p.seBni + c];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = min(m0 + 32 * wm + OZ32_ROW(lg, r), p.M - 1);
        const int ea = p.eA[b * p.seAb + (int64_t)m * p.seAm + c];
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int d = 0; d < NS; ++d) {  // exact integers times 254^-d
          sr = fma((double)ar[d][r], oz_weight(d), sr);
          si = fma((double)ai[d][r], oz_weight(d), si);
        }
        const double sc = ldexp(1.0 / (OZ_FIRST * OZ_FIRST), ea + eb);
        cr[r] = fma(sr, sc, cr[r]);
        ci[r] = fma(si, sc, ci[r]);
      }
#pragma unroll
      for (int d = 0; d < NS; ++d)
#pragma unroll
        for (int r = 0; r < 16; ++r) { ar[d][r] = 0; ai[d][r] = 0; }
    }
  };

  load_stage(oz_s0, 0);
  __syncthreads();
  for (int g = 0; g < nsteps; g += 2) {
    if (g + 1 < nsteps) load_stage(oz_s1, g + 1);
    compute(oz_s0, g);
    __syncthreads();
    if (g + 1 >= nsteps) break;
    if (g + 2 < nsteps) load_stage(oz_s0, g + 2);
    compute(oz_s1, g + 1);
    __syncthreads();
  }

  const int n = n0 + 32 * wn + l32;
  if (n >= p.N) return;
  const int no = n / p.nin;
  const int64_t coff = (int64_t)no * p.sCno + (int64_t)(n - no * p.nin) * p.sCni;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + 32 * wm + OZ32_ROW(lg, r);
    if (m >= p.M) continue;
    const double vr = cr[r], vi = ci[r];
    double2 v = make_double2(p.alpha.x * vr - p.alpha.y * vi, p.alpha.x * vi + p.alpha.y * vr);
    double2* dst = p.C + b * p.sCb + (int64_t)m * p.sCm + coff;
    if (p.accumulate) {
      const double2 o = *dst;
      v.x += o.x;
      v.y += o.y;
    }
    *dst = v;
  }
}

// About 64 blocks (the ones resident on one XCD) share a group when there are few column
// tiles; with many column tiles 16 row tiles of A stay in L2.
inline int ozaki_default_group(int tiles_n) { return tiles_n >= 64 ? 16 : std::max(1, 64 / tiles_n); }

template <typename K>
inline void oz_allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) GWHIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

inline size_t ozaki_pipe_lds_bytes(int ns) { return (size_t)2 * OZ_PLANES * OZ_BM * ns * 64; }

template <int NS>
inline void ozaki_launch(const OzGemmArgs& p, int64_t grid, hipStream_t stream) {
  if (p.tile == 4) {
    hipLaunchKernelGGL(zgemm_ozaki32_kernel<NS>, dim3((unsigned)grid), dim3(256), 0, stream, p);
    return;
  }
  const size_t lds = ozaki_pipe_lds_bytes(NS);
  if (p.nchunk == 1) {
    oz_allow_lds(zgemm_ozaki_pipe_kernel<NS, true>, lds);
    hipLaunchKernelGGL((zgemm_ozaki_pipe_kernel<NS, true>), dim3((unsigned)grid), dim3(256), lds, stream, p);
  } else {
    oz_allow_lds(zgemm_ozaki_pipe_kernel<NS, false>, lds);
    hipLaunchKernelGGL((zgemm_ozaki_pipe_kernel<NS, false>), dim3((unsigned)grid), dim3(256), lds, stream, p);
  }
}

inline bool ozaki_tile_valid(int tile) { return tile == 2 || tile == 4; }

inline void zgemm_ozaki(OzGemmArgs p, int ns, hipStream_t stream) {
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return;
  if (p.kc % OZ_KS != 0 || p.nchunk <= 0 || p.nin <= 0) throw std::runtime_error("zgemm_ozaki: bad chunking");
  if (p.kc > split_kc_limit(ns)) throw std::runtime_error("zgemm_ozaki: chunk too long for exact int32 sums");
  if (p.tile == 0) p.tile = 2;
  if (!ozaki_tile_valid(p.tile)) throw std::runtime_error("zgemm_ozaki: unknown block tile");
  p.tiles_m = (int)ceil_div(p.M, p.tile == 4 ? OZ32_BM : OZ_BM);
  p.tiles_n = (int)ceil_div(p.N, p.tile == 4 ? OZ32_BN : OZ_BN);
  p.nblocks = (int64_t)p.batch * p.tiles_m * p.tiles_n;
  if (p.group_m <= 0) p.group_m = ozaki_default_group(p.tiles_n);
  p.group_m = (int)std::min<int64_t>(p.group_m, (int64_t)p.batch * p.tiles_m);
  p.per_xcd = ceil_div(p.nblocks, 8);
  const int64_t grid = p.per_xcd * 8;
  if (grid > 0x7fffffffLL) throw std::runtime_error("zgemm_ozaki: grid too large");
  if (ns == 4) ozaki_launch<4>(p, grid, stream);
  else if (ns == 5) ozaki_launch<5>(p, grid, stream);
  else if (ns == 3) ozaki_launch<3>(p, grid, stream);
  else throw std::runtime_error("zgemm_ozaki: slices must be 3, 4 or 5");
  GWHIP_CHECK(hipGetLastError());
}

inline void slice_rows(const SliceArgs& a, int ns, hipStream_t stream) {
  const int64_t tasks = a.nrows * a.nchunk;
  if (tasks <= 0) return;
  if (a.kc % 64 != 0 || a.klen > a.kc) throw std::runtime_error("slice_rows: bad chunk length");
  const dim3 g((unsigned)ceil_div(tasks, 4));
  if (ns == 4) hipLaunchKernelGGL(slice_rows_kernel<4>, g, dim3(256), 0, stream, a);
  else if (ns == 5) hipLaunchKernelGGL(slice_rows_kernel<5>, g, dim3(256), 0, stream, a);
  else if (ns == 3) hipLaunchKernelGGL(slice_rows_kernel<3>, g, dim3(256), 0, stream, a);
  else throw std::runtime_error("slice_rows: slices must be 3, 4 or 5");
  GWHIP_CHECK(hipGetLastError());
}

}  // namespace gwhip
