// Split-precision complex GEMM: FP64 operands cut into int8 slices, products on the int8
// matrix cores (v_mfma_i32_16x16x64_i8) with exact int32 accumulation, FP64 accumulation
// across K chunks.  This is the fixed-point slicing of Ozaki et al. (Numer. Algorithms 59,
// 2012; Ootomo, Ozaki & Yokota, IJHPCA 2024 for integer matrix units):
//
//   x = 2^e * sum_{s<NS} x_s 2^{-7(s+1)} + r,   x_s in [-127, 127],  |r| < 2^{e - 7 NS}
//
// with one power-of-two exponent e per (row, K chunk) of A and per (column, K chunk) of B,
// so every slice product is an integer and sum_k x_s y_t is exact in int32 while
// 2^14 * (chunk length) * (number of slice pairs) < 2^31.  Slice pairs with s + t = d share
// the weight 2^{-7(d+2)}; pairs with s + t >= NS are dropped.
//
// Slice store (int8), for each row (A) or column (B):
//   [chunk][plane][slice][kc]   planes Re, Im
// (B is stored already conjugated when the contraction needs conj(B)), so
//   Re C = Re A Re B + (-Im A) Im B,   Im C = Re A Im B + Im A Re B,
// with -Im A negated in registers (slices lie in [-127, 127], so negation is exact).
#pragma once

#include <algorithm>

#include "common.hpp"

namespace gwhip {

typedef int v4i __attribute__((ext_vector_type(4)));

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
  int nplanes;   // 2: Re, Im   (3: also -Im; not used by the GEMMs)
  int8_t* dst;   // [row][chunk][plane][slice][kc]
  int* exps;     // [row][chunk]
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
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
  const double sc = ldexp(1.0, -e);
  const double sgn = a.conj ? -1.0 : 1.0;
  int8_t* out = a.dst + task * (int64_t)a.nplanes * NS * a.kc;
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
      double tr = v.x * sc, ti = sgn * v.y * sc;  // |t| < 1
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        tr *= 128.0;
        ti *= 128.0;
        const double qr = trunc(tr), qi = trunc(ti);
        tr -= qr;
        ti -= qi;
        wr[t] |= ((int)qr & 0xff) << (8 * j);
        wi[t] |= ((int)qi & 0xff) << (8 * j);
      }
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      *reinterpret_cast<int*>(out + (int64_t)(0 * NS + t) * a.kc + k0) = wr[t];
      *reinterpret_cast<int*>(out + (int64_t)(1 * NS + t) * a.kc + k0) = wi[t];
      if (a.nplanes == 3) {
        int nw = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int8_t bq = (int8_t)((wi[t] >> (8 * j)) & 0xff);
          nw |= ((int)(int8_t)(-bq) & 0xff) << (8 * j);
        }
        *reinterpret_cast<int*>(out + (int64_t)(2 * NS + t) * a.kc + k0) = nw;
      }
    }
  }
}

// ------------------------------------------------------------------------------ GEMM
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

// Largest K chunk (bytes per plane-slice, a multiple of 64) whose int32 accumulation stays exact
// with ns slices: an accumulator of slice level d sums up to ns slice pairs of two planes, each
// product below 2^14 in magnitude, so 2 * ns * 2^14 * kc < 2^31.
inline int64_t split_kc_limit(int ns) { return ((((int64_t)1 << 31) - 1) / ((int64_t)2 * ns << 14)) / 64 * 64; }

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
  int tile;     // 0: 32 x 64 block tile, two blocks per CU; 1: 64 x 64 with register prefetch, one per CU;
                // 2: 32 x 64, software-pipelined (next K step in flight, double-buffered A)
  int upper;    // 1: skip block tiles that lie wholly below the diagonal (Hermitian products)
};

// Output layout of v_mfma_i32_16x16x64_i8: register r of lane l holds row
// OZ_ROW(l/16, r), column l%16 (checked by the identity-probe test).
#define OZ_ROW(g, r) (4 * (g) + (r))

// Block tile 32 x 64 (complex); the four waves split N, each owning 32 x 16 = two 16x16
// blocks.  A (32 rows, shared by all waves) is staged through LDS one K step (64 bytes per
// plane-slice row) at a time, its four 16-byte pieces XOR-swizzled by (r >> 2) & 3 so that
// the 16 rows of one fragment read fall on distinct banks.  Each wave reads its own 16
// B columns straight from global memory into MFMA fragments (no other wave uses them).
// SQ_LDS_BANK_CONFLICT per LDS instruction: 4.8 cycles when each LDS store of A covered two
// plane-slices 2 KB apart, 3.2 with the store of 16 whole rows of one plane-slice (a 1 % faster
// GEMM); a second swizzle, distinct banks also for every 8 rows of equal parity, left the count
// unchanged.  A layout in MFMA fragment order (piece-major within 16 rows: every fragment read a
// contiguous 1 KB) had no conflicts at all but ran slower: Sigma GEMM 987 ms per evaluation with
// row-fastest staging loads, 950 ms with piece-fastest ones, against 929 ms here.
constexpr int OZ_BM = 32, OZ_BN = 64, OZ_KS = 64;

__device__ __forceinline__ int oz_lds_off(int ps, int rows, int r, int piece) {
  return (ps * rows + r) * 64 + 16 * (piece ^ ((r >> 2) & 3));
}

template <int NS>
__global__ __launch_bounds__(256, 2) void zgemm_ozaki_kernel(OzGemmArgs p) {
  constexpr int PA = 2;
  extern __shared__ __attribute__((aligned(16))) int8_t oz_lds[];
  int8_t* As = oz_lds;                                          // [PA*NS][32 rows][64 B]
  int* eAs = reinterpret_cast<int*>(As + PA * NS * OZ_BM * 64);  // [32]
  int* eBs = eAs + OZ_BM;                                        // [64]

  // Each XCD takes a contiguous range of logical tiles; within it, groups of group_m row tiles
  // (rows run over batch x M) sweep every column tile, row fastest, so a group's A panels stay in
  // that XCD's L2 while the B panels stream through it once per group.
  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return;
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

  // A staging slots: 32 rows x 4 pieces, two threads per slot splitting the plane-slices
  // A staging: thread -> (plane-slice parity, row, 16-byte piece) with the piece fastest, so the
  // 64 lanes of one LDS store cover 16 whole rows of one plane-slice (1 KB, no bank conflicts)
  const int a_row = (tid >> 2) & 31, a_piece = tid & 3, a_par = tid >> 7;
  const int8_t* a_ptr = nullptr;
  if (m0 + a_row < p.M) a_ptr = p.A + b * p.sAb + (int64_t)(m0 + a_row) * p.sAm + 16 * a_piece;
  const int a_lds = oz_lds_off(0, OZ_BM, a_row, a_piece);
  // this lane's B column and K-step piece
  const int bcol = n0 + wn + lrow;
  const int8_t* b_ptr = nullptr;
  if (bcol < p.N) {
    const int no = bcol / p.nin;
    b_ptr = p.B + b * p.sBb + (int64_t)no * p.sBno + (int64_t)(bcol - no * p.nin) * p.sBni + 16 * lg;
  }

  double cr[2][4], ci[2][4];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int r = 0; r < 4; ++r) { cr[x][r] = 0.0; ci[x][r] = 0.0; }

  for (int c = 0; c < p.nchunk; ++c) {
    v4i ar[NS][2], ai[NS][2];  // [d][x] int32 accumulators (Re, Im)
#pragma unroll
    for (int d = 0; d < NS; ++d)
#pragma unroll
      for (int x = 0; x < 2; ++x) { ar[d][x] = v4i{0, 0, 0, 0}; ai[d][x] = v4i{0, 0, 0, 0}; }
    const int8_t* a_c = a_ptr ? a_ptr + (int64_t)c * p.sAc : nullptr;
    const int8_t* b_c = b_ptr ? b_ptr + (int64_t)c * p.sBc : nullptr;

    for (int ks = 0; ks < p.kc; ks += OZ_KS) {
      __syncthreads();
#pragma unroll
      for (int ps = a_par; ps < PA * NS; ps += 2) {
        v4i v = v4i{0, 0, 0, 0};
        if (a_c) v = *reinterpret_cast<const v4i*>(a_c + (int64_t)ps * p.kc + ks);
        *reinterpret_cast<v4i*>(As + a_lds + ps * OZ_BM * 64) = v;
      }
      // B fragments of this K step, all slices (Re, Im)
      v4i fb_r[NS], fb_i[NS];
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        fb_r[t] = v4i{0, 0, 0, 0};
        fb_i[t] = v4i{0, 0, 0, 0};
        if (b_c) {
          fb_r[t] = *reinterpret_cast<const v4i*>(b_c + (int64_t)(0 * NS + t) * p.kc + ks);
          fb_i[t] = *reinterpret_cast<const v4i*>(b_c + (int64_t)(1 * NS + t) * p.kc + ks);
        }
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        v4i fa_r[2], fa_i[2], fa_n[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          const int row = 16 * x + lrow;
          fa_r[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(0 * NS + s, OZ_BM, row, lg));
          fa_i[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(1 * NS + s, OZ_BM, row, lg));
          fa_n[x] = neg_i8x16(fa_i[x]);
        }
#pragma unroll
        for (int t = 0; t + s < NS; ++t) {
#pragma unroll
          for (int x = 0; x < 2; ++x) {
            ar[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb_r[t], ar[s + t][x], 0, 0, 0);
            ai[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb_i[t], ai[s + t][x], 0, 0, 0);
            ar[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_n[x], fb_i[t], ar[s + t][x], 0, 0, 0);
            ai[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_i[x], fb_r[t], ai[s + t][x], 0, 0, 0);
          }
        }
      }
    }
    // chunk end: int32 -> FP64 with the (row, column, chunk) power-of-two scale
    __syncthreads();
    if (tid < OZ_BM) {
      const int m = m0 + tid;
      eAs[tid] = (m < p.M) ? p.eA[b * p.seAb + (int64_t)m * p.seAm + c] : 0;
    } else if (tid < OZ_BM + OZ_BN) {
      const int n = n0 + tid - OZ_BM;
      int e = 0;
      if (n < p.N) {
        const int no = n / p.nin;
        e = p.eB[b * p.seBb + (int64_t)no * p.seBno + (int64_t)(n - no * p.nin) * p.seBni + c];
      }
      eBs[tid - OZ_BM] = e;
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * x + OZ_ROW(lg, r), col = wn + lrow;
        double sc = ldexp(1.0, eAs[row] + eBs[col] - 14);
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int d = 0; d < NS; ++d) {
          sr = fma((double)ar[d][x][r], sc, sr);
          si = fma((double)ai[d][x][r], sc, si);
          sc *= 0.0078125;  // 2^-7
        }
        cr[x][r] += sr;
        ci[x][r] += si;
      }
  }

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

// The 32 x 64 tile with a software pipeline: while the MFMAs of K step g run, the A pieces and B
// fragments of step g + 1 are already in flight into registers; A goes to the other half of a
// double-buffered LDS panel after the multiplies, so one barrier per K step suffices and no
// wave waits on global memory between steps.  The chunk exponents are read straight from
// global memory (L2) in the chunk epilogue.  Same products, same order: results are bit
// identical to zgemm_ozaki_kernel.
template <int NS>
__global__ __launch_bounds__(256, 2) void zgemm_ozaki_pipe_kernel(OzGemmArgs p) {
  constexpr int PA = 2, NPC = PA * NS / 2;  // plane-slices staged per thread and K step
  constexpr int BUF = PA * NS * OZ_BM * 64;
  extern __shared__ __attribute__((aligned(16))) int8_t oz_lds[];

  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return;
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

  // A staging: thread -> (plane-slice parity, row, 16-byte piece) with the piece fastest, so the
  // 64 lanes of one LDS store cover 16 whole rows of one plane-slice (1 KB, no bank conflicts)
  const int a_row = (tid >> 2) & 31, a_piece = tid & 3, a_par = tid >> 7;
  const int8_t* a_ptr = nullptr;
  if (m0 + a_row < p.M) a_ptr = p.A + b * p.sAb + (int64_t)(m0 + a_row) * p.sAm + 16 * a_piece;
  const int a_lds = oz_lds_off(0, OZ_BM, a_row, a_piece);
  const int bcol = n0 + wn + lrow;
  const int8_t* b_ptr = nullptr;
  if (bcol < p.N) {
    const int no = bcol / p.nin;
    b_ptr = p.B + b * p.sBb + (int64_t)no * p.sBno + (int64_t)(bcol - no * p.nin) * p.sBni + 16 * lg;
  }
  const int nks = p.kc / OZ_KS, nsteps = p.nchunk * nks;
  auto b_at = [&](int g) -> const int8_t* {  // this lane's B bytes of K step g
    const int c = g / nks, ks = (g - c * nks) * OZ_KS;
    return b_ptr + (int64_t)c * p.sBc + ks;
  };

  // A pieces of K step g into registers, later stored to one half of the LDS panel
  v4i a_nx[NPC];
  auto load_a = [&](int g) {
    const int c = g / nks, ks = (g - c * nks) * OZ_KS;
#pragma unroll
    for (int j = 0; j < NPC; ++j) {
      const int ps = a_par + 2 * j;
      a_nx[j] = a_ptr ? *reinterpret_cast<const v4i*>(a_ptr + (int64_t)c * p.sAc + (int64_t)ps * p.kc + ks)
                      : v4i{0, 0, 0, 0};
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int j = 0; j < NPC; ++j)
      *reinterpret_cast<v4i*>(oz_lds + buf * BUF + a_lds + (a_par + 2 * j) * OZ_BM * 64) = a_nx[j];
  };

  double cr[2][4], ci[2][4];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int r = 0; r < 4; ++r) { cr[x][r] = 0.0; ci[x][r] = 0.0; }
  v4i ar[NS][2], ai[NS][2];
#pragma unroll
  for (int d = 0; d < NS; ++d)
#pragma unroll
    for (int x = 0; x < 2; ++x) { ar[d][x] = v4i{0, 0, 0, 0}; ai[d][x] = v4i{0, 0, 0, 0}; }

  // B fragments of the current step; fb[t] is last used by slice s = NS - 1 - t, after which its
  // registers take the next step's fragment, so B for step g + 1 is in flight during step g
  // without a second register set
  v4i fb_r[NS], fb_i[NS];
  load_a(0);
  store_a(0);
  {
    const int8_t* bc = b_ptr ? b_at(0) : nullptr;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      fb_r[t] = bc ? *reinterpret_cast<const v4i*>(bc + (int64_t)(0 * NS + t) * p.kc) : v4i{0, 0, 0, 0};
      fb_i[t] = bc ? *reinterpret_cast<const v4i*>(bc + (int64_t)(1 * NS + t) * p.kc) : v4i{0, 0, 0, 0};
    }
  }
  __syncthreads();

  for (int g = 0; g < nsteps; ++g) {
    const bool more = g + 1 < nsteps;
    const int c = g / nks;
    const bool chunk_end = g - c * nks == nks - 1;
    if (more) load_a(g + 1);
    const int8_t* bn = (more && b_ptr) ? b_at(g + 1) : nullptr;
    const int8_t* As = oz_lds + (g & 1) * BUF;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      v4i fa_r[2], fa_i[2], fa_n[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int row = 16 * x + lrow;
        fa_r[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(0 * NS + s, OZ_BM, row, lg));
        fa_i[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(1 * NS + s, OZ_BM, row, lg));
        fa_n[x] = neg_i8x16(fa_i[x]);
      }
#pragma unroll
      for (int t = 0; t + s < NS; ++t) {
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          ar[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb_r[t], ar[s + t][x], 0, 0, 0);
          ai[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb_i[t], ai[s + t][x], 0, 0, 0);
          ar[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_n[x], fb_i[t], ar[s + t][x], 0, 0, 0);
          ai[s + t][x] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_i[x], fb_r[t], ai[s + t][x], 0, 0, 0);
        }
      }
      if (more) {
        const int t = NS - 1 - s;
        fb_r[t] = bn ? *reinterpret_cast<const v4i*>(bn + (int64_t)(0 * NS + t) * p.kc) : v4i{0, 0, 0, 0};
        fb_i[t] = bn ? *reinterpret_cast<const v4i*>(bn + (int64_t)(1 * NS + t) * p.kc) : v4i{0, 0, 0, 0};
      }
    }
    if (chunk_end) {  // int32 -> FP64 with the (row, column, chunk) scale
      int eb = 0;
      const int n = n0 + wn + lrow;
      if (n < p.N) {
        const int no = n / p.nin;
        eb = p.eB[b * p.seBb + (int64_t)no * p.seBno + (int64_t)(n - no * p.nin) * p.seBni + c];
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + 16 * x + OZ_ROW(lg, r);
          const int ea = m < p.M ? p.eA[b * p.seAb + (int64_t)m * p.seAm + c] : 0;
          double sc = ldexp(1.0, ea + eb - 14);
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int d = 0; d < NS; ++d) {
            sr = fma((double)ar[d][x][r], sc, sr);
            si = fma((double)ai[d][x][r], sc, si);
            sc *= 0.0078125;  // 2^-7
          }
          cr[x][r] += sr;
          ci[x][r] += si;
        }
#pragma unroll
      for (int d = 0; d < NS; ++d)
#pragma unroll
        for (int x = 0; x < 2; ++x) { ar[d][x] = v4i{0, 0, 0, 0}; ai[d][x] = v4i{0, 0, 0, 0}; }
    }
    if (more) store_a((g + 1) & 1);
    __syncthreads();
  }

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

// About 64 blocks (the ones resident on one XCD) share a group when there are few column
// tiles; with many column tiles 16 row tiles of A (~2.5 MB at K = 320, 5 slices) stay in L2.
inline int ozaki_default_group(int tiles_n) { return tiles_n >= 64 ? 16 : std::max(1, 64 / tiles_n); }

// 64 x 64 block tile, 2 x 2 waves of 32 x 32, one block per CU.  A (15 plane-slices x 64 rows)
// and B (10 plane-slices x 64 columns) of one K step sit in LDS; the next K step is loaded into
// registers while the current one is multiplied (240 MFMAs per wave per K step hide the L2
// latency without a second resident block).  The int32 accumulators are meant for AGPRs.
constexpr int OZ4_BM = 64, OZ4_BN = 64;

__device__ __forceinline__ void oz64_store(const OzGemmArgs& p, int64_t b, int m, int n, double vr, double vi) {
  if (m >= p.M || n >= p.N) return;
  const int no = n / p.nin;
  double2 v = make_double2(p.alpha.x * vr - p.alpha.y * vi, p.alpha.x * vi + p.alpha.y * vr);
  double2* dst = p.C + b * p.sCb + (int64_t)m * p.sCm + (int64_t)no * p.sCno + (int64_t)(n - no * p.nin) * p.sCni;
  if (p.accumulate) {
    const double2 o = *dst;
    v.x += o.x;
    v.y += o.y;
  }
  *dst = v;
}

template <int NS, bool MULTI>
__global__ __launch_bounds__(256, 1) void zgemm_ozaki64_kernel(OzGemmArgs p) {
  constexpr int PA = 2, PB = 2;
  extern __shared__ __attribute__((aligned(16))) int8_t oz_lds[];
  int8_t* As = oz_lds;                                         // [PA*NS][64 rows][64 B]
  int8_t* Bs = As + PA * NS * OZ4_BM * 64;                       // [PB*NS][64 cols][64 B]
  int* eAs = reinterpret_cast<int*>(Bs + PB * NS * OZ4_BN * 64);  // [64]
  int* eBs = eAs + OZ4_BM;                                       // [64]

  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw % 8) * p.per_xcd + hw / 8;
  if (logical >= p.nblocks) return;
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
  const int m0 = tm * OZ4_BM, n0 = tn * OZ4_BN;
  if (p.upper && m0 >= n0 + OZ4_BN) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lg = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // staging: thread -> (row / column, 16-byte piece); all plane-slices of that piece
  const int s_row = tid >> 2, s_piece = tid & 3;
  const int8_t* a_ptr = nullptr;
  if (m0 + s_row < p.M) a_ptr = p.A + b * p.sAb + (int64_t)(m0 + s_row) * p.sAm + 16 * s_piece;
  const int8_t* b_ptr = nullptr;
  if (n0 + s_row < p.N) {
    const int col = n0 + s_row, no = col / p.nin;
    b_ptr = p.B + b * p.sBb + (int64_t)no * p.sBno + (int64_t)(col - no * p.nin) * p.sBni + 16 * s_piece;
  }
  const int a_lds = oz_lds_off(0, OZ4_BM, s_row, s_piece);
  const int b_lds = oz_lds_off(0, OZ4_BN, s_row, s_piece);

  constexpr int NC = MULTI ? 2 : 1;  // FP64 accumulators across chunks only when there are several
  double cr[NC][2][4], ci[NC][2][4];
#pragma unroll
  for (int x = 0; x < NC; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) { cr[x][y][r] = 0.0; ci[x][y][r] = 0.0; }

  v4i ra[PA * NS], rb[PB * NS];
  auto load = [&](int c, int ks) {
#pragma unroll
    for (int ps = 0; ps < PA * NS; ++ps)
      ra[ps] = a_ptr ? *reinterpret_cast<const v4i*>(a_ptr + (int64_t)c * p.sAc + (int64_t)ps * p.kc + ks) : v4i{0, 0, 0, 0};
#pragma unroll
    for (int ps = 0; ps < PB * NS; ++ps)
      rb[ps] = b_ptr ? *reinterpret_cast<const v4i*>(b_ptr + (int64_t)c * p.sBc + (int64_t)ps * p.kc + ks) : v4i{0, 0, 0, 0};
  };
  auto store = [&]() {
#pragma unroll
    for (int ps = 0; ps < PA * NS; ++ps) *reinterpret_cast<v4i*>(As + a_lds + ps * OZ4_BM * 64) = ra[ps];
#pragma unroll
    for (int ps = 0; ps < PB * NS; ++ps) *reinterpret_cast<v4i*>(Bs + b_lds + ps * OZ4_BN * 64) = rb[ps];
  };

  const int steps = p.kc / OZ_KS;
  load(0, 0);
  for (int c = 0; c < p.nchunk; ++c) {
    v4i ar[NS][2][2], ai[NS][2][2];
#pragma unroll
    for (int d = 0; d < NS; ++d)
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) { ar[d][x][y] = v4i{0, 0, 0, 0}; ai[d][x][y] = v4i{0, 0, 0, 0}; }
    for (int st = 0; st < steps; ++st) {
      __syncthreads();
      store();
      __syncthreads();
      // prefetch the next K step (next chunk after the last step of this one)
      if (st + 1 < steps) load(c, (st + 1) * OZ_KS);
      else if (c + 1 < p.nchunk) load(c + 1, 0);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        v4i fa_r[2], fa_i[2], fa_n[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          const int row = wm + 16 * x + lrow;
          fa_r[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(0 * NS + s, OZ4_BM, row, lg));
          fa_i[x] = *reinterpret_cast<const v4i*>(As + oz_lds_off(1 * NS + s, OZ4_BM, row, lg));
          fa_n[x] = neg_i8x16(fa_i[x]);
        }
#pragma unroll
        for (int t = 0; t + s < NS; ++t) {
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            const int col = wn + 16 * y + lrow;
            const v4i fb_r = *reinterpret_cast<const v4i*>(Bs + oz_lds_off(0 * NS + t, OZ4_BN, col, lg));
            const v4i fb_i = *reinterpret_cast<const v4i*>(Bs + oz_lds_off(1 * NS + t, OZ4_BN, col, lg));
#pragma unroll
            for (int x = 0; x < 2; ++x) {
              ar[s + t][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb_r, ar[s + t][x][y], 0, 0, 0);
              ai[s + t][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_r[x], fb_i, ai[s + t][x][y], 0, 0, 0);
              ar[s + t][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_n[x], fb_i, ar[s + t][x][y], 0, 0, 0);
              ai[s + t][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa_i[x], fb_r, ai[s + t][x][y], 0, 0, 0);
            }
          }
        }
        // keep the scheduler from hoisting every slice's fragment loads (register pressure)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // chunk end: int32 -> FP64 with the (row, column, chunk) power-of-two scale
    if (tid < OZ4_BM) {
      const int m = m0 + tid;
      eAs[tid] = (m < p.M) ? p.eA[b * p.seAb + (int64_t)m * p.seAm + c] : 0;
    } else if (tid < OZ4_BM + OZ4_BN) {
      const int n = n0 + tid - OZ4_BM;
      int e = 0;
      if (n < p.N) {
        const int no = n / p.nin;
        e = p.eB[b * p.seBb + (int64_t)no * p.seBno + (int64_t)(n - no * p.nin) * p.seBni + c];
      }
      eBs[tid - OZ4_BM] = e;
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm + 16 * x + OZ_ROW(lg, r), col = wn + 16 * y + lrow;
          double sc = ldexp(1.0, eAs[row] + eBs[col] - 14);
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int d = 0; d < NS; ++d) {
            sr = fma((double)ar[d][x][y][r], sc, sr);
            si = fma((double)ai[d][x][y][r], sc, si);
            sc *= 0.0078125;  // 2^-7
          }
          if (MULTI) {
            cr[x % NC][y][r] += sr;  // NC == 2 here
            ci[x % NC][y][r] += si;
          } else {
            oz64_store(p, b, m0 + row, n0 + col, sr, si);
          }
        }
    __syncthreads();  // eAs / eBs are rewritten by the next chunk
  }
  if (MULTI) {
#pragma unroll
    for (int x = 0; x < NC; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          oz64_store(p, b, m0 + wm + 16 * x + OZ_ROW(lg, r), n0 + wn + 16 * y + lrow, cr[x][y][r], ci[x][y][r]);
  }
}

inline size_t ozaki_lds_bytes(int ns) {
  return (size_t)2 * OZ_BM * ns * 64 + (OZ_BM + OZ_BN) * sizeof(int);
}

template <typename K>
inline void oz_allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) GWHIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

inline size_t ozaki_pipe_lds_bytes(int ns) { return (size_t)2 * 2 * OZ_BM * ns * 64; }

inline size_t ozaki64_lds_bytes(int ns) {
  return (size_t)(2 * OZ4_BM + 2 * OZ4_BN) * ns * 64 + (OZ4_BM + OZ4_BN) * sizeof(int);
}

template <int NS>
inline void ozaki_launch(const OzGemmArgs& p, int64_t grid, hipStream_t stream) {
  if (p.tile == 1 && p.nchunk > 1) {
    const size_t lds = ozaki64_lds_bytes(NS);
    oz_allow_lds(zgemm_ozaki64_kernel<NS, true>, lds);
    hipLaunchKernelGGL((zgemm_ozaki64_kernel<NS, true>), dim3((unsigned)grid), dim3(256), lds, stream, p);
  } else if (p.tile == 1) {
    const size_t lds = ozaki64_lds_bytes(NS);
    oz_allow_lds(zgemm_ozaki64_kernel<NS, false>, lds);
    hipLaunchKernelGGL((zgemm_ozaki64_kernel<NS, false>), dim3((unsigned)grid), dim3(256), lds, stream, p);
  } else if (p.tile == 2) {
    const size_t lds = ozaki_pipe_lds_bytes(NS);
    oz_allow_lds(zgemm_ozaki_pipe_kernel<NS>, lds);
    hipLaunchKernelGGL(zgemm_ozaki_pipe_kernel<NS>, dim3((unsigned)grid), dim3(256), lds, stream, p);
  } else {
    const size_t lds = ozaki_lds_bytes(NS);
    oz_allow_lds(zgemm_ozaki_kernel<NS>, lds);
    hipLaunchKernelGGL(zgemm_ozaki_kernel<NS>, dim3((unsigned)grid), dim3(256), lds, stream, p);
  }
}

inline void zgemm_ozaki(OzGemmArgs p, int ns, hipStream_t stream) {
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return;
  if (p.kc % OZ_KS != 0 || p.nchunk <= 0 || p.nin <= 0) throw std::runtime_error("zgemm_ozaki: bad chunking");
  if (p.tile < 0 || p.tile > 2) throw std::runtime_error("zgemm_ozaki: tile must be 0, 1 or 2");
  // exact int32 accumulation: ns pairs at most share one accumulator, 4 products each ... bounded
  if ((double)p.kc * 2.0 * 127.0 * 127.0 * ns > 2147483647.0) throw std::runtime_error("zgemm_ozaki: chunk too long");
  const int BM = p.tile == 1 ? OZ4_BM : OZ_BM, BN = p.tile == 1 ? OZ4_BN : OZ_BN;
  p.tiles_m = (int)ceil_div(p.M, BM);
  p.tiles_n = (int)ceil_div(p.N, BN);
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
