// Orbital-size (nao^3 * NQ) contractions of the GW path on FP64 matrix cores.
//
// Both kernels chain two v_mfma_f64_16x16x4f64 products without an LDS round trip: for a
// 16x16 result block the MFMA leaves row (l/16) + 4r, column l%16 in register r of lane l,
// which is exactly the operand layout a following MFMA expects for its K-slab r (A operand:
// row l%16, k l/16; B operand: column l%16, k l/16).  Orbital dimensions are padded to
// NB blocks of 16 (nao <= 16 NB); complex products are four real MFMAs.
//
// Layouts (complex = double2):
//   V [k][Q][p][m]  = V^Q_pm(k, k +- q)     the resident integrals of one q (the only copy)
//   Z [tz][k][Q][t][n]                       P0 left operand (row-major per Q)
//   R [tz][k][Q][n][j]                       P-tilde conj(V), per-Q transposed
// The orbital products take V_Q as the operand whose K index runs over its second orbital index,
// so a lane group reads four consecutive elements of a row (64 bytes) per slab.
#pragma once

#include "../../include/gwhip.h"
#include "common.hpp"
#include "gw_kernels.hpp"
#include "zgemm_ozaki.hpp"

namespace gwhip {

// Block order of the orbital kernels.  Consecutive hardware workgroup ids go to different XCDs,
// so each XCD is given a contiguous range of logical blocks, and the logical index runs over tau
// fastest: the blocks that read the same V(k, .) panel for different tau then run on one XCD
// and find it in that XCD's L2.  Logical (x, y, z) over a gx x gy x gz grid: z fastest, then x.
struct XcdGrid {
  int gx, gy, gz;
  int64_t per_xcd;  // ceil(gx gy gz / 8); the launch has 8 per_xcd workgroups
};

__device__ __forceinline__ bool xcd_block(const XcdGrid& g, int& x, int& y, int& z) {
  const int64_t hw = blockIdx.x;
  const int64_t l = (hw % 8) * g.per_xcd + hw / 8;
  if (l >= (int64_t)g.gx * g.gy * g.gz) return false;
  z = (int)(l % g.gz);
  const int64_t r = l / g.gz;
  x = (int)(r % g.gx);
  y = (int)(r / g.gx);
  return true;
}

// These kernels run 256 work-items per work-group, and a launch holds under 2^32 work-items.
inline XcdGrid make_xcd_grid(dim3 g) {
  XcdGrid x{(int)g.x, (int)g.y, (int)g.z, 0};
  x.per_xcd = ((int64_t)g.x * g.y * g.z + 7) / 8;
  if (x.per_xcd * 8 * 256 > (int64_t)0xffffffffLL)
    throw status_error(GWHIP_ERR_UNSUPPORTED, "orbital kernel launch exceeds 2^32 work-items");
  return x;
}
inline dim3 xcd_launch_grid(const XcdGrid& x) { return dim3((unsigned)(x.per_xcd * 8)); }

typedef float v4f __attribute__((ext_vector_type(4)));

// Matrix-core flavour of the orbital contractions: FP64 (v_mfma_f64_16x16x4f64) for the
// plain path, FP32 (v_mfma_f32_16x16x4f32) for the split-precision path.  acc_t is the
// 4-register accumulator; row_of(g, r) is the result row held in register r by lane group g.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row_of(int g, int r) { return g + 4 * r; }
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
};
template <> struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row_of(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
};

// (cr, ci) += (ar + i ai) * (br + i bi) on 16x16x4 fragments
template <typename T>
__device__ __forceinline__ void cmfma(typename Mfma<T>::acc_t& cr, typename Mfma<T>::acc_t& ci, T ar, T ai, T br, T bi) {
  cr = Mfma<T>::mma(ar, br, cr);
  ci = Mfma<T>::mma(ar, bi, ci);
  cr = Mfma<T>::mma(-ai, bi, cr);
  ci = Mfma<T>::mma(ai, br, ci);
}
__device__ __forceinline__ void cmfma(v4d& cr, v4d& ci, double ar, double ai, double br, double bi) {
  cmfma<double>(cr, ci, ar, ai, br, bi);
}

__device__ __forceinline__ v4d v4zero() { return v4d{0.0, 0.0, 0.0, 0.0}; }

// Complex accumulator of a 16x16 MFMA fragment chain: four real MFMAs per complex product
// (Re += ar br - ai bi, Im += ar bi + ai br).  Gauss's three-product form (p1 = ar br, p2 = ai bi,
// p3 = (ar + ai)(br + bi)) was measured slower in these kernels: Sigma epilogue 1522 ms against
// 1100 ms per evaluation at nk = 64, nao = 26, naux = 300, P0 prep 643 against 571 ms.
template <typename T> struct CAcc {
  typename Mfma<T>::acc_t cr, ci;
  __device__ __forceinline__ void zero() { cr = ci = Mfma<T>::zero(); }
  __device__ __forceinline__ void mac(T ar, T ai, T br, T bi) { cmfma<T>(cr, ci, ar, ai, br, bi); }
  __device__ __forceinline__ T re(int r) const { return cr[r]; }
  __device__ __forceinline__ T im(int r) const { return ci[r]; }
};

// ---------------------------------------------------------------------------------------
// P0 left operand  Z_Q = sum_s Gb_s V_Q Gq_s  (Gb_s = G(beta - tau, k) block (b, a),
// Gq_s = G(tau, k + q) block (a, b)); src/gw_cpu_kernel.cpp:170-213 regrouped.
// grid: (ceil(NQ / (4 QPW)), nk, tcount); 4 waves, each handles QPW consecutive Q.
// LDS per spin block: Gb and Gq as zero-padded real/imag planes [16NB][16NB+1] of T.
// T = double: FP64 matrix cores; T = float: FP32 matrix cores (split-precision path), K <= 16NB
// per MFMA chain, result written in FP64.  Registers: at most half the file for NB <= 2 (two or
// more waves per SIMD), all of it for NB = 3.
template <int NB, typename T>
__global__ __launch_bounds__(256, NB < 3 ? 2 : 1) void z_prep_mfma_kernel(XcdGrid xg, DevDims d, const double2* __restrict__ G,
                                                          const double2* __restrict__ V,
                                                          double2* __restrict__ Z, const int* __restrict__ kplus,
                                                          int t0, int QPW) {
  typedef Mfma<T> M;
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  T* Gbr = reinterpret_cast<T*>(lds_d);
  T* Gbi = Gbr + PL;
  T* Gqr = Gbi + PL;
  T* Gqi = Gqr + PL;
  const int nao = d.nao, nao2 = nao * nao;
  int bx, by, bz;
  if (!xcd_block(xg, bx, by, bz)) return;
  const int tz = bz, tau = t0 + tz, tt = d.nts - 1 - tau;
  const int k = by, kq = kplus[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int nslab = (nao + 3) / 4;

  for (int s = 0; s < d.pseudo_ns; ++s) {
    const int is = s / (d.pns * d.pns), a = (s / d.pns) % d.pns, b = s % d.pns;
    const double2* Gbar = G + g_index(d, tt, is, k);
    const double2* Gk = G + g_index(d, tau, is, kq);
    __syncthreads();
    for (int e = tid; e < N16 * N16; e += blockDim.x) {
      const int r = e / N16, c = e - (e / N16) * N16;
      double2 gb = make_double2(0.0, 0.0), gq = make_double2(0.0, 0.0);
      if (r < nao && c < nao) {
        gb = Gbar[(int64_t)(b * nao + r) * d.nso + a * nao + c];  // Gb[t = r][p = c]
        gq = Gk[(int64_t)(a * nao + r) * d.nso + b * nao + c];    // Gq[m = r][n = c]
      }
      Gbr[r * LD + c] = (T)gb.x;
      Gbi[r * LD + c] = (T)gb.y;
      Gqr[r * LD + c] = (T)gq.x;
      Gqi[r * LD + c] = (T)gq.y;
    }
    __syncthreads();
    for (int qi = 0; qi < QPW; ++qi) {
      const int Q = (bx * 4 + wave) * QPW + qi;
      if (Q >= d.NQ) break;
      const double2* vq = V + ((int64_t)k * d.NQ + Q) * nao2;  // vq[p * nao + m] = V[p][m]
      double2* zq = Z + (((int64_t)tz * d.nk + k) * d.NQ + Q) * nao2;
      // W[:, bn] = V Gq[:, bn] for every column block, then Z[:, bn] (+)= Gb W[:, bn] for all of them
      // at once (each Gb element read from LDS serves every column block; NB^2 independent
      // accumulator pairs); the same sums in the same order as column block by column block
      CAcc<T> w[NB][NB];  // [bn][bp]
#pragma unroll
      for (int bn = 0; bn < NB; ++bn)
#pragma unroll
        for (int x = 0; x < NB; ++x) w[bn][x].zero();
      for (int sl = 0; sl < nslab; ++sl) {
        const int m = 4 * sl + lk;
        double2 av[NB];
#pragma unroll
        for (int bp = 0; bp < NB; ++bp) {
          const int p = 16 * bp + lrow;
          av[bp] = (p < nao && m < nao) ? vq[p * nao + m] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int bn = 0; bn < NB; ++bn) {
          const T br = Gqr[m * LD + 16 * bn + lrow], bi = Gqi[m * LD + 16 * bn + lrow];
#pragma unroll
          for (int bp = 0; bp < NB; ++bp) w[bn][bp].mac((T)av[bp].x, (T)av[bp].y, br, bi);
        }
      }
      // register r of W holds row p = 16 bp + row_of(lk, r)
      CAcc<T> z[NB][NB];  // [bn][bt]
#pragma unroll
      for (int bn = 0; bn < NB; ++bn)
#pragma unroll
        for (int x = 0; x < NB; ++x) z[bn][x].zero();
#pragma unroll
      for (int bp = 0; bp < NB; ++bp)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (sizeof(T) == 8 && 16 * bp + 4 * r >= nao) continue;  // K slab wholly in the padding
          const int p = 16 * bp + M::row_of(lk, r);
#pragma unroll
          for (int bt = 0; bt < NB; ++bt) {
            const int t = 16 * bt + lrow;
            const T gr = Gbr[t * LD + p], gi = Gbi[t * LD + p];
#pragma unroll
            for (int bn = 0; bn < NB; ++bn) z[bn][bt].mac(gr, gi, w[bn][bp].re(r), w[bn][bp].im(r));
          }
        }
#pragma unroll
      for (int bn = 0; bn < NB; ++bn) {
        const int n = 16 * bn + lrow;
        if (n >= nao) continue;
#pragma unroll
        for (int bt = 0; bt < NB; ++bt)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int t = 16 * bt + M::row_of(lk, r);
            if (t < nao) {
              double2 v = make_double2((double)z[bn][bt].re(r), (double)z[bn][bt].im(r));
              if (s > 0) {
                const double2 o = zq[t * nao + n];
                v.x += o.x;
                v.y += o.y;
              }
              zq[t * nao + n] = v;
            }
          }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Split-precision P0 operand straight into int8 slices (one or two spin blocks, pseudo_ns <= 2;
// the G planes of every block sit in LDS and Z_Q = sum_s Gb_s V_Q Gq_s accumulates over s in the
// same registers): the wave that forms Z_Q(tau, k) in registers also owns the whole slice row-chunk
// (row (tau, Q), chunk k, nao^2 elements), so it finds the chunk exponent with a wave reduction
// and writes the NS slices of both planes (Re, Im) itself -- the same arithmetic as
// slice_rows_kernel, so the slices are bit-identical to slicing the FP64 Z, without Z ever reaching
// HBM.  Slice layout: dst[((tz*NQ + Q)*nk + k)][plane][slice][kc], exps[(tz*NQ + Q)*nk + k].
// Two waves per SIMD for NB <= 2 (three spill and run 15 % slower at nao = 26).
template <int NB, int NS, int SB>
__global__ __launch_bounds__(256, NB < 3 ? 2 : 1) void z_prep_slice_kernel(XcdGrid xg, DevDims d, const double2* __restrict__ G,
                                                           const double2* __restrict__ V, int8_t* __restrict__ dst,
                                                           int* __restrict__ exps, int kc,
                                                           const int* __restrict__ kplus, int t0, int QPW) {
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  double* Gbr = lds_d;
  double* Gbi = Gbr + PL;
  double* Gqr = Gbi + PL;
  double* Gqi = Gqr + PL;
  const int nao = d.nao, nao2 = nao * nao;
  int bx, by, bz;
  if (!xcd_block(xg, bx, by, bz)) return;
  const int tz = bz, tau = t0 + tz, tt = d.nts - 1 - tau;
  const int k = by, kq = kplus[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int nslab = (nao + 3) / 4;
  for (int s = 0; s < SB; ++s) {  // SB = pseudo_ns collinear spin blocks: is = s, a = b = 0
    const double2* Gbar = G + g_index(d, tt, s, k);
    const double2* Gk = G + g_index(d, tau, s, kq);
    for (int e = tid; e < N16 * N16; e += blockDim.x) {
      const int r = e / N16, c = e - (e / N16) * N16;
      double2 gb = make_double2(0.0, 0.0), gq = make_double2(0.0, 0.0);
      if (r < nao && c < nao) {
        gb = Gbar[(int64_t)r * d.nso + c];
        gq = Gk[(int64_t)r * d.nso + c];
      }
      Gbr[s * 4 * PL + r * LD + c] = gb.x;
      Gbi[s * 4 * PL + r * LD + c] = gb.y;
      Gqr[s * 4 * PL + r * LD + c] = gq.x;
      Gqi[s * 4 * PL + r * LD + c] = gq.y;
    }
  }
  __syncthreads();
  for (int qi = 0; qi < QPW; ++qi) {
    const int Q = (bx * 4 + wave) * QPW + qi;
    if (Q >= d.NQ) break;
    const double2* vq = V + ((int64_t)k * d.NQ + Q) * nao2;  // vq[p * nao + m] = V[p][m]
    CAcc<double> z[NB][NB];                                 // [bn][bt]
    double2 avs[4 * NB][NB];                                    // V_Q, loaded once per Q
#pragma unroll
    for (int sl = 0; sl < 4 * NB; ++sl)
#pragma unroll
      for (int bp = 0; bp < NB; ++bp) {
        const int m = 4 * sl + lk, p = 16 * bp + lrow;
        avs[sl][bp] = (p < nao && m < nao) ? vq[p * nao + m] : make_double2(0.0, 0.0);
      }
#pragma unroll
    for (int bn = 0; bn < NB; ++bn)
#pragma unroll
      for (int x = 0; x < NB; ++x) z[bn][x].zero();
#pragma unroll
    for (int s = 0; s < SB; ++s) {
      const double* gbr = Gbr + s * 4 * PL;
      const double* gbi = Gbi + s * 4 * PL;
      const double* gqr = Gqr + s * 4 * PL;
      const double* gqi = Gqi + s * 4 * PL;
      // W[:, bn] = V_Q Gq[:, bn] for every column block, then Z[:, bn] += Gb W[:, bn] for all of
      // them at once: each Gb element read from LDS serves every column block and the second
      // product has NB^2 independent accumulator pairs; the same sums in the same order as block
      // by block
      CAcc<double> w[NB][NB];  // [bn][bp]
#pragma unroll
      for (int bn = 0; bn < NB; ++bn) {
#pragma unroll
        for (int x = 0; x < NB; ++x) w[bn][x].zero();
#pragma unroll
        for (int sl = 0; sl < 4 * NB; ++sl) {
          if (NB < 3 && sl >= nslab) break;
          if (sl < nslab) {  // NB = 3: a guard, not a break, so that the loop unrolls and avs stays in registers
            const int m = 4 * sl + lk;
            const double br = gqr[m * LD + 16 * bn + lrow], bi = gqi[m * LD + 16 * bn + lrow];
#pragma unroll
            for (int bp = 0; bp < NB; ++bp) w[bn][bp].mac(avs[sl][bp].x, avs[sl][bp].y, br, bi);
          }
        }
      }
#pragma unroll
      for (int bp = 0; bp < NB; ++bp)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (16 * bp + 4 * r >= nao) continue;  // K slab wholly in the padding
          const int p = 16 * bp + lk + 4 * r;
#pragma unroll
          for (int bt = 0; bt < NB; ++bt) {
            const int t = 16 * bt + lrow;
            const double gr = gbr[t * LD + p], gi = gbi[t * LD + p];
#pragma unroll
            for (int bn = 0; bn < NB; ++bn) z[bn][bt].mac(gr, gi, w[bn][bp].re(r), w[bn][bp].im(r));
          }
        }
    }
    // Re and Im of the finished products
    v4d zr[NB][NB], zi[NB][NB];
#pragma unroll
    for (int bn = 0; bn < NB; ++bn)
#pragma unroll
      for (int bt = 0; bt < NB; ++bt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          zr[bn][bt][r] = z[bn][bt].re(r);
          zi[bn][bt][r] = z[bn][bt].im(r);
        }
    // chunk exponent: max |Re|, |Im| over the nao x nao tile (padding entries are exact zeros)
    double mx = 0.0;
#pragma unroll
    for (int bn = 0; bn < NB; ++bn)
#pragma unroll
      for (int bt = 0; bt < NB; ++bt)
#pragma unroll
        for (int r = 0; r < 4; ++r) mx = fmax(mx, fmax(fabs(zr[bn][bt][r]), fabs(zi[bn][bt][r])));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    const int64_t task = ((int64_t)tz * d.NQ + Q) * d.nk + k;
    if (lane == 0) exps[task] = e;
    const double sc = ldexp(OZ_FIRST, -e);
    int8_t* out = dst + task * (int64_t)OZ_PLANES * NS * kc;
#pragma unroll
    for (int bn = 0; bn < NB; ++bn) {
      const int n = 16 * bn + lrow;
#pragma unroll
      for (int bt = 0; bt < NB; ++bt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = 16 * bt + lk + 4 * r;
          if (n >= nao || t >= nao) continue;
          const int kk = t * nao + n;
          int dr[NS], di[NS];
          oz_element_digits<NS>(zr[bn][bt][r], zi[bn][bt][r], sc, dr, di);
#pragma unroll
          for (int sl = 0; sl < NS; ++sl) {
            out[(int64_t)(0 * NS + sl) * kc + kk] = (int8_t)dr[sl];
            out[(int64_t)(1 * NS + sl) * kc + kk] = (int8_t)di[sl];
          }
        }
    }
    // zero the chunk padding [nao^2, kc) of every plane-slice
    const int padn = kc - nao2;
    for (int i = lane; i < OZ_PLANES * NS * padn; i += 64) {
      const int ps = i / padn, kk = nao2 + (i - ps * padn);
      out[(int64_t)ps * kc + kk] = 0;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Sigma(tau, is, k)[a,b](i, j) += scale * sum_Q sum_n T_Q[i][n] R_Q[n][j],  T_Q = V_Q G_s
// (upper != 0: 16 x 16 blocks below the diagonal are skipped; Sigma is Hermitian for Hermitian
// G and the caller mirrors it once all q are in)
// with V = V(k, k - q), G_s = G(tau, k - q) block (a, b)  (src/gw_cpu_kernel.cpp:359-373,
// Y2 = Y1^T P done by the big GEMM as R = P-tilde conj(V)).  The first MFMA produces
// T^T = G^T V^T so that its result registers feed the second MFMA's A operand.
// grid: (nks, qsplit, tcount) in the XcdGrid order (tcount = 2 x the R rows when mirror is set);
// block y takes the y-th of qsplit equal Q ranges, wave w its Q = w, w + 4, ...; the four partial
// sums meet in LDS.  qsplit = 1 (part = nullptr) adds to Sigma; qsplit > 1 stores each range's sum
// to part[y] ([qsplit][nts][ns][nks][nso^2]), which epi_part_reduce_kernel then adds to Sigma in
// range order (more work-groups when nks x tcount alone would not fill the chip).  Block x = k is the position in this
// launch's k list: V (V[k] = V(k, k - q)), R ([t][nks][Q][nao^2]) and Sigma ([nts][ns][nks][nso^2]) use it
// directly, G (full mesh) is read at kminus[k].  The full mesh has nks = nk; the star-of-q
// evaluation passes the irreducible k points only.
// With T = float each Q's product (K <= 16 NB) is added to FP64 accumulators.
// Occupancy: for NB <= 2 the register budget is set for three waves per SIMD (168 VGPRs and a
// 112-byte spill per lane; 825 ms per evaluation at nk = 64, nao = 26, naux = 300 against 958 ms at
// two waves without a spill).  V_Q is loaded slab by slab inside the first product.
template <int NB, typename T>
__global__ __launch_bounds__(256, NB < 3 ? 3 : 1) void sigma_epi_mfma_kernel(XcdGrid xg, DevDims d, const double2* __restrict__ G,
                                                             const double2* __restrict__ V,
                                                             const double2* __restrict__ R,
                                                             double2* __restrict__ Sigma,
                                                             const int* __restrict__ kminus, int nks, int t0,
                                                             double scale, int upper, int mirror, int qsplit,
                                                             double2* __restrict__ part) {
  typedef Mfma<T> M;
  constexpr bool F64 = sizeof(T) == 8;
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  T* Gr = reinterpret_cast<T*>(lds_d);
  T* Gi = Gr + PL;
  double* red = lds_d + PL;  // [NB*NB*8][64], after the T planes (2 PL T <= PL doubles)
  const int nao = d.nao, nao2 = nao * nao;
  int bx, by, bz;
  if (!xcd_block(xg, bx, by, bz)) return;
  // mirror != 0: R holds tau = t0 + bz/2 of the half grid and serves tau and beta - tau
  // (P-tilde(beta - tau) = P-tilde(tau)); the pair runs on neighbouring blocks
  const int tz = mirror ? (bz >> 1) : bz;
  const int k = bx, tau = mirror && (bz & 1) ? d.nts - 1 - (t0 + tz) : t0 + tz;
  const int kmq = kminus[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int nslab = (nao + 3) / 4;
  const int qspan = (d.NQ + qsplit - 1) / qsplit, qlo = by * qspan, qhi = min(d.NQ, qlo + qspan);

  for (int s = 0; s < d.pseudo_ns; ++s) {
    const int is = s / (d.pns * d.pns), a = (s / d.pns) % d.pns, b = s % d.pns;
    const double2* Gk = G + g_index(d, tau, is, kmq);
    __syncthreads();
    for (int e = tid; e < N16 * N16; e += blockDim.x) {
      const int r = e / N16, c = e - (e / N16) * N16;
      double2 g = make_double2(0.0, 0.0);
      if (r < nao && c < nao) g = Gk[(int64_t)(a * nao + r) * d.nso + b * nao + c];  // G[m = r][n = c]
      Gr[r * LD + c] = (T)g.x;
      Gi[r * LD + c] = (T)g.y;
    }
    __syncthreads();
    double sr[NB][NB][4], si[NB][NB][4];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) { sr[x][y][r] = 0.0; si[x][y][r] = 0.0; }
    CAcc<T> acc64[NB][NB];  // FP64 flavour accumulates straight in the MFMA registers
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y) acc64[x][y].zero();

    for (int Q = qlo + wave; Q < qhi; Q += 4) {
      const double2* vq = V + ((int64_t)k * d.NQ + Q) * nao2;                    // vq[i * nao + m] = V[i][m]
      const double2* rq = R + (((int64_t)tz * nks + k) * d.NQ + Q) * nao2;      // rq[n * nao + j]
      if (F64) {
        // all row blocks at once: each G element read from LDS and each R element loaded once per Q
        // serves every row block, and the second product feeds up to NB (NB + 1) / 2 independent
        // accumulator pairs (one at a time, row block bi = 1 of NB = 2 had a single pair, a
        // dependent MFMA every other issue); the same sums in the same order as block by block
        CAcc<T> tq[NB][NB];  // [bi][bn]: T^T block (bn, bi)
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
          for (int y = 0; y < NB; ++y) tq[x][y].zero();
#pragma unroll
        for (int sl = 0; sl < 4 * NB; ++sl) {
          if (sl >= nslab) break;
          const int m = 4 * sl + lk;
          double2 bv[NB];
#pragma unroll
          for (int bi = 0; bi < NB; ++bi) {
            const int i = 16 * bi + lrow;
            bv[bi] = (m < nao && i < nao) ? vq[i * nao + m] : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int bn = 0; bn < NB; ++bn) {
            const int n = 16 * bn + lrow;
            const T gr = Gr[m * LD + n], gi = Gi[m * LD + n];
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) tq[bi][bn].mac(gr, gi, (T)bv[bi].x, (T)bv[bi].y);
          }
        }
#pragma unroll
        for (int bn = 0; bn < NB; ++bn)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (16 * bn + 4 * r >= nao) continue;  // K slab wholly in the padding
            const int n = 16 * bn + M::row_of(lk, r);
            double2 rv[NB];
#pragma unroll
            for (int bj = 0; bj < NB; ++bj) {
              const int j = 16 * bj + lrow;
              rv[bj] = (n < nao && j < nao) ? rq[n * nao + j] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) {
              const T xr = tq[bi][bn].re(r), xi = tq[bi][bn].im(r);
#pragma unroll
              for (int bj = 0; bj < NB; ++bj) {
                if (upper && bj < bi) continue;  // block below the diagonal: mirrored afterwards
                acc64[bi][bj].mac(xr, xi, (T)rv[bj].x, (T)rv[bj].y);
              }
            }
          }
        continue;
      }
#pragma unroll
      for (int bi = 0; bi < NB; ++bi) {
        CAcc<T> tq[NB];
#pragma unroll
        for (int x = 0; x < NB; ++x) tq[x].zero();
        const int i = 16 * bi + lrow;
#pragma unroll
        for (int sl = 0; sl < 4 * NB; ++sl) {
          if (sl >= nslab) break;
          const int m = 4 * sl + lk;
          const double2 bv = (m < nao && i < nao) ? vq[i * nao + m] : make_double2(0.0, 0.0);
#pragma unroll
          for (int bn = 0; bn < NB; ++bn) {
            const int n = 16 * bn + lrow;
            tq[bn].mac(Gr[m * LD + n], Gi[m * LD + n], (T)bv.x, (T)bv.y);
          }
        }
        CAcc<T> pq[NB];
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          if (F64) pq[y] = acc64[bi][y];
          else pq[y].zero();
        }
#pragma unroll
        for (int bn = 0; bn < NB; ++bn)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (sizeof(T) == 8 && 16 * bn + 4 * r >= nao) continue;  // K slab wholly in the padding
            const int n = 16 * bn + M::row_of(lk, r);
            const T xr = tq[bn].re(r), xi = tq[bn].im(r);
#pragma unroll
            for (int bj = 0; bj < NB; ++bj) {
              if (upper && bj < bi) continue;  // block below the diagonal: mirrored afterwards
              const int j = 16 * bj + lrow;
              double2 rv = make_double2(0.0, 0.0);
              if (n < nao && j < nao) rv = rq[n * nao + j];
              pq[bj].mac(xr, xi, (T)rv.x, (T)rv.y);
            }
          }
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          if (F64) {
            acc64[bi][y] = pq[y];
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              sr[bi][y][r] += (double)pq[y].re(r);
              si[bi][y][r] += (double)pq[y].im(r);
            }
          }
        }
      }
    }
    if (F64) {
#pragma unroll
      for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            sr[x][y][r] = (double)acc64[x][y].re(r);
            si[x][y][r] = (double)acc64[x][y].im(r);
          }
    }
    // reduce the four waves' partial sums into wave 0
    for (int w = 1; w < 4; ++w) {
      __syncthreads();
      if (wave == w) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
          for (int y = 0; y < NB; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              red[(((x * NB + y) * 2 + 0) * 4 + r) * 64 + lane] = sr[x][y][r];
              red[(((x * NB + y) * 2 + 1) * 4 + r) * 64 + lane] = si[x][y][r];
            }
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
          for (int y = 0; y < NB; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              sr[x][y][r] += red[(((x * NB + y) * 2 + 0) * 4 + r) * 64 + lane];
              si[x][y][r] += red[(((x * NB + y) * 2 + 1) * 4 + r) * 64 + lane];
            }
      }
    }
    if (wave == 0) {
      // whole Q range: added to Sigma; one of qsplit ranges: stored to its own partial slab
      const int64_t off = (((int64_t)tau * d.ns + is) * nks + k) * d.nso * d.nso;
      double2* S = part ? part + (int64_t)by * d.nts * d.ns * nks * d.nso * d.nso + off : Sigma + off;
#pragma unroll
      for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          const int j = 16 * y + lrow;
          if (j >= nao) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * x + M::row_of(lk, r);
            if (i >= nao) continue;
            double2* dst = S + (int64_t)(a * nao + i) * d.nso + b * nao + j;
            double2 v = part ? make_double2(0.0, 0.0) : *dst;
            v.x += scale * sr[x][y][r];
            v.y += scale * si[x][y][r];
            *dst = v;
          }
        }
    }
  }
}

// ---------------------------------------------------------------------------------------
// The Sigma epilogue above on v_mfma_f64_4x4x4_4b, for nao <= 4 NB4 <= 32.  One instruction does four
// independent 4 x 4 x 4 products; its lane layouts (checked by the layout-probe test) are
//   A_b[r][k] at lane 16 k + 4 b + r,   B_b[k][c] at lane 16 k + 4 b + c,   D_b[r][c] at lane 16 r + 4 b + c,
// and it issues at the FP64 matrix-core rate with only four independent accumulators per wave.  The four
// products of an instruction are four consecutive 4-wide blocks of i (a group of 16 i), so orbital
// dimensions pad to multiples of 4 (n, m, j) and 16 (i) instead of 16 everywhere:
//   T^T[n][i] = sum_m G[m][n] V[i][m]           A = G^T block (the same for the four), B = V^T
//   Sigma^T[j][i] += sum_n R[n][j] T^T[n][i]    A = R^T block (the same for the four), B = T^T
// Lane L holds T^T[n0 + L/16][16 ig + L%16] after the first product of K slab n0, which is exactly
// the B operand of the second product for that slab, so the intermediate stays in registers.  V_Q
// (the first product's B operands, all slabs) is loaded once per Q; R_Q one K slab of n ahead.
// Grid, Q ranges, partial sums and the upper-triangle skipping as in sigma_epi_mfma_kernel: a tile
// (j block, i group) is skipped when every j in it lies below every i.
__device__ __forceinline__ void cmfma4(double& cr, double& ci, double ar, double ai, double br, double bi) {
  cr = __builtin_amdgcn_mfma_f64_4x4x4f64(ar, br, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f64_4x4x4f64(ar, bi, ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f64_4x4x4f64(-ai, bi, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f64_4x4x4f64(ai, br, ci, 0, 0, 0);
}

template <int NB4>
__global__ __launch_bounds__(256, NB4 <= 4 ? 3 : 2) void sigma_epi4_kernel(XcdGrid xg, DevDims d, const double2* __restrict__ G,
                                                            const double2* __restrict__ V,
                                                            const double2* __restrict__ R,
                                                            double2* __restrict__ Sigma,
                                                            const int* __restrict__ kminus, int nks, int t0,
                                                            double scale, int upper, int mirror, int qsplit,
                                                            double2* __restrict__ part) {
  constexpr int NG = (NB4 + 3) / 4;          // groups of 16 i
  constexpr int N4 = 4 * NB4, LD = N4 + 1;   // G planes [N4][LD] in LDS
  extern __shared__ double lds_d[];
  double* Gr = lds_d;
  double* Gi = Gr + N4 * LD;
  double* red = Gi + N4 * LD;  // [NB4 * NG][2][64]
  const int nao = d.nao, nao2 = nao * nao;
  int bx, by, bz;
  if (!xcd_block(xg, bx, by, bz)) return;
  const int tz = mirror ? (bz >> 1) : bz;
  const int k = bx, tau = mirror && (bz & 1) ? d.nts - 1 - (t0 + tz) : t0 + tz;
  const int kmq = kminus[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lk = lane >> 4, l16 = lane & 15, l4 = lane & 3;
  const int qspan = (d.NQ + qsplit - 1) / qsplit, qlo = by * qspan, qhi = min(d.NQ, qlo + qspan);

  for (int s = 0; s < d.pseudo_ns; ++s) {
    const int is = s / (d.pns * d.pns), a = (s / d.pns) % d.pns, b = s % d.pns;
    const double2* Gk = G + g_index(d, tau, is, kmq);
    __syncthreads();
    for (int e = tid; e < N4 * N4; e += blockDim.x) {
      const int r = e / N4, c = e - (e / N4) * N4;
      double2 g = make_double2(0.0, 0.0);
      if (r < nao && c < nao) g = Gk[(int64_t)(a * nao + r) * d.nso + b * nao + c];  // G[m = r][n = c]
      Gr[r * LD + c] = g.x;
      Gi[r * LD + c] = g.y;
    }
    __syncthreads();
    double sr[NB4][NG], si[NB4][NG];  // Sigma^T tile (j block, i group), this wave's Q
#pragma unroll
    for (int x = 0; x < NB4; ++x)
#pragma unroll
      for (int y = 0; y < NG; ++y) { sr[x][y] = 0.0; si[x][y] = 0.0; }

    for (int Q = qlo + wave; Q < qhi; Q += 4) {
      const double2* vq = V + ((int64_t)k * d.NQ + Q) * nao2;                // vq[i * nao + m] = V[i][m]
      const double2* rq = R + (((int64_t)tz * nks + k) * d.NQ + Q) * nao2;  // rq[n * nao + j]
      double vr[NB4][NG], vi[NB4][NG];  // B of the first product: V[16 ig + L%16][4 ms + L/16]
#pragma unroll
      for (int ms = 0; ms < NB4; ++ms)
#pragma unroll
        for (int ig = 0; ig < NG; ++ig) {
          const int i = 16 * ig + l16, m = 4 * ms + lk;
          const double2 v = (i < nao && m < nao) ? vq[i * nao + m] : make_double2(0.0, 0.0);
          vr[ms][ig] = v.x;
          vi[ms][ig] = v.y;
        }
#pragma unroll
      for (int nb = 0; nb < NB4; ++nb) {
        if (4 * nb >= nao) break;  // K slab wholly in the padding
        const int n = 4 * nb + lk;
        double rr[NB4], ri[NB4];  // A of the second product: R[4 nb + L/16][4 jb + L%4]
#pragma unroll
        for (int jb = 0; jb < NB4; ++jb) {
          const int j = 4 * jb + l4;
          const double2 r = (n < nao && j < nao) ? rq[n * nao + j] : make_double2(0.0, 0.0);
          rr[jb] = r.x;
          ri[jb] = r.y;
        }
        double tr[NG], ti[NG];
#pragma unroll
        for (int ig = 0; ig < NG; ++ig) { tr[ig] = 0.0; ti[ig] = 0.0; }
#pragma unroll
        for (int ms = 0; ms < NB4; ++ms) {
          if (4 * ms >= nao) break;
          const int m = 4 * ms + lk, nn = 4 * nb + l4;  // A = G^T[n][m]: G[4 ms + L/16][4 nb + L%4]
          const double gr = Gr[m * LD + nn], gi = Gi[m * LD + nn];
#pragma unroll
          for (int ig = 0; ig < NG; ++ig) cmfma4(tr[ig], ti[ig], gr, gi, vr[ms][ig], vi[ms][ig]);
        }
#pragma unroll
        for (int jb = 0; jb < NB4; ++jb)
#pragma unroll
          for (int ig = 0; ig < NG; ++ig) {
            if (upper && 4 * jb + 3 < 16 * ig) continue;  // every j below every i: mirrored afterwards
            cmfma4(sr[jb][ig], si[jb][ig], rr[jb], ri[jb], tr[ig], ti[ig]);
          }
      }
    }
    // reduce the four waves' partial sums into wave 0
    for (int w = 1; w < 4; ++w) {
      __syncthreads();
      if (wave == w) {
#pragma unroll
        for (int x = 0; x < NB4; ++x)
#pragma unroll
          for (int y = 0; y < NG; ++y) {
            red[((x * NG + y) * 2 + 0) * 64 + lane] = sr[x][y];
            red[((x * NG + y) * 2 + 1) * 64 + lane] = si[x][y];
          }
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int x = 0; x < NB4; ++x)
#pragma unroll
          for (int y = 0; y < NG; ++y) {
            sr[x][y] += red[((x * NG + y) * 2 + 0) * 64 + lane];
            si[x][y] += red[((x * NG + y) * 2 + 1) * 64 + lane];
          }
      }
    }
    if (wave == 0) {
      const int64_t off = (((int64_t)tau * d.ns + is) * nks + k) * d.nso * d.nso;
      double2* S = part ? part + (int64_t)by * d.nts * d.ns * nks * d.nso * d.nso + off : Sigma + off;
#pragma unroll
      for (int x = 0; x < NB4; ++x)
#pragma unroll
        for (int y = 0; y < NG; ++y) {
          const int j = 4 * x + lk, i = 16 * y + l16;  // lane L: Sigma^T[4 x + L/16][16 y + L%16]
          if (i >= nao || j >= nao) continue;
          if (upper && 4 * x + 3 < 16 * y) continue;
          double2* dst = S + (int64_t)(a * nao + i) * d.nso + b * nao + j;
          double2 v = part ? make_double2(0.0, 0.0) : *dst;
          v.x += scale * sr[x][y];
          v.y += scale * si[x][y];
          *dst = v;
        }
    }
  }
}

inline size_t sigma_epi4_lds_bytes(int nb4) {
  const int n4 = 4 * nb4, ng = (nb4 + 3) / 4;
  return ((size_t)2 * n4 * (n4 + 1) + (size_t)nb4 * ng * 2 * 64) * sizeof(double);
}

// Buffer-resource word 3 on CDNA (gfx9 family): 32-bit data format, no swizzle; a raw (stride 0)
// resource returns zero for a load at or past its byte size.
constexpr int BUF_RSRC_WORD3 = 0x00020000;
typedef double v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v2d buf_load_f64x2(__amdgpu_buffer_rsrc_t r, int byte_offset) {
  return __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(r, byte_offset, 0, 0));
}

// The same products with one group of 16 i per wave: a work-group of NG x QW waves, wave w owning
// i group w % NG and every QW-th Q of the range from w / NG on.  The columns i of T^T and Sigma^T
// are independent, so splitting them over waves costs no MFMA work; each wave holds V_Q, R_Q
// and the Sigma^T tile of its own i group only (12 NB4 + 4 VGPRs of operands and accumulators),
// which keeps nao up to 48 (NB4 = 12, three i groups) at three waves per SIMD where the all-groups
// form above runs out of registers.  A wave loads only the R blocks its upper-triangle tiles
// need.  The QW partial sums of one i group are added in the order of w / NG.
template <int NB4, int QW, int WPE>
__global__ __launch_bounds__(64 * ((NB4 + 3) / 4) * QW) __attribute__((amdgpu_waves_per_eu(WPE)))
void sigma_epi4w_kernel(XcdGrid xg, DevDims d, const double2* __restrict__ G, const double2* __restrict__ V,
                        const double2* __restrict__ R, double2* __restrict__ Sigma,
                        const int* __restrict__ kminus, int nks, int t0, double scale, int upper, int mirror,
                        int qsplit, double2* __restrict__ part) {
  constexpr int NG = (NB4 + 3) / 4;
  constexpr int N4 = 4 * NB4, LD = N4 + 1;
  constexpr int NT = 64 * NG * QW;
  extern __shared__ double lds_d[];
  double* Gr = lds_d;
  double* Gi = Gr + N4 * LD;
  double* red = Gi + N4 * LD;  // [QW - 1][NG][NB4][2][64]
  const int nao = d.nao, nao2 = nao * nao;
  int bx, by, bz;
  if (!xcd_block(xg, bx, by, bz)) return;
  const int tz = mirror ? (bz >> 1) : bz;
  const int k = bx, tau = mirror && (bz & 1) ? d.nts - 1 - (t0 + tz) : t0 + tz;
  const int kmq = kminus[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ig = wave % NG, qw = wave / NG;
  const int lk = lane >> 4, l16 = lane & 15, l4 = lane & 3;
  const int i = 16 * ig + l16;
  const int jmin = upper ? 4 * ig : 0;  // j blocks below it lie wholly below every i of the group
  const int qspan = (d.NQ + qsplit - 1) / qsplit, qlo = by * qspan, qhi = min(d.NQ, qlo + qspan);

  for (int s = 0; s < d.pseudo_ns; ++s) {
    const int is = s / (d.pns * d.pns), a = (s / d.pns) % d.pns, b = s % d.pns;
    const double2* Gk = G + g_index(d, tau, is, kmq);
    __syncthreads();
    for (int e = tid; e < N4 * N4; e += NT) {
      const int r = e / N4, c = e - (e / N4) * N4;
      double2 g = make_double2(0.0, 0.0);
      if (r < nao && c < nao) g = Gk[(int64_t)(a * nao + r) * d.nso + b * nao + c];  // G[m = r][n = c]
      Gr[r * LD + c] = g.x;
      Gi[r * LD + c] = g.y;
    }
    __syncthreads();
    double sr[NB4], si[NB4];  // Sigma^T[4 jb + L/16][16 ig + L%16]
#pragma unroll
    for (int x = 0; x < NB4; ++x) { sr[x] = 0.0; si[x] = 0.0; }

    for (int Q = qlo + qw; Q < qhi; Q += QW) {
      const double2* vq = V + ((int64_t)k * d.NQ + Q) * nao2;                // vq[i * nao + m] = V[i][m]
      const double2* rq = R + (((int64_t)tz * nks + k) * d.NQ + Q) * nao2;  // rq[n * nao + j]
      // V_Q and R_Q through buffer resources bounded to their nao^2 elements: a row past nao
      // reads zero, a column past nao the next row's first elements (finite).  A padded i or j
      // only reaches Sigma^T entries that are never stored, and a padded m or n meets the zero
      // padding of G in the same product; the per-slab offsets are immediates.
      const __amdgpu_buffer_rsrc_t vres = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<double2*>(vq), (short)0, nao2 * (int)sizeof(double2), BUF_RSRC_WORD3);
      const __amdgpu_buffer_rsrc_t rres = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<double2*>(rq), (short)0, nao2 * (int)sizeof(double2), BUF_RSRC_WORD3);
      double vr[NB4], vi[NB4];  // B of the first product: V[16 ig + L%16][4 ms + L/16]
      const int voff = (i * nao + lk) * (int)sizeof(double2);
#pragma unroll
      for (int ms = 0; ms < NB4; ++ms) {
        const v2d v = buf_load_f64x2(vres, voff + ms * 4 * (int)sizeof(double2));
        vr[ms] = v.x;
        vi[ms] = v.y;
      }
#pragma unroll 1
      for (int nb = 0; nb < NB4; ++nb) {
        if (4 * nb >= nao) break;  // K slab wholly in the padding
        const int roff = ((4 * nb + lk) * nao + l4) * (int)sizeof(double2);
        double rr[NB4], ri[NB4];  // A of the second product: R[4 nb + L/16][4 jb + L%4]
#pragma unroll
        for (int jb = 0; jb < NB4; ++jb) {
          if (jb < jmin || 4 * jb >= nao) continue;
          const v2d r = buf_load_f64x2(rres, roff + jb * 4 * (int)sizeof(double2));
          rr[jb] = r.x;
          ri[jb] = r.y;
        }
        double tr = 0.0, ti = 0.0;
#pragma unroll
        for (int ms = 0; ms < NB4; ++ms) {
          if (4 * ms >= nao) break;
          const int m = 4 * ms + lk, nn = 4 * nb + l4;  // A = G^T[n][m]: G[4 ms + L/16][4 nb + L%4]
          cmfma4(tr, ti, Gr[m * LD + nn], Gi[m * LD + nn], vr[ms], vi[ms]);
        }
#pragma unroll
        for (int jb = 0; jb < NB4; ++jb) {
          if (jb < jmin || 4 * jb >= nao) continue;
          cmfma4(sr[jb], si[jb], rr[jb], ri[jb], tr, ti);
        }
      }
    }
    if (QW > 1) {
      if (qw > 0) {
#pragma unroll
        for (int x = 0; x < NB4; ++x) {
          red[((((qw - 1) * NG + ig) * NB4 + x) * 2 + 0) * 64 + lane] = sr[x];
          red[((((qw - 1) * NG + ig) * NB4 + x) * 2 + 1) * 64 + lane] = si[x];
        }
      }
      __syncthreads();
      if (qw == 0) {
        for (int w = 1; w < QW; ++w)
#pragma unroll
          for (int x = 0; x < NB4; ++x) {
            sr[x] += red[((((w - 1) * NG + ig) * NB4 + x) * 2 + 0) * 64 + lane];
            si[x] += red[((((w - 1) * NG + ig) * NB4 + x) * 2 + 1) * 64 + lane];
          }
      }
    }
    if (qw == 0 && i < nao) {
      const int64_t off = (((int64_t)tau * d.ns + is) * nks + k) * d.nso * d.nso;
      double2* S = part ? part + (int64_t)by * d.nts * d.ns * nks * d.nso * d.nso + off : Sigma + off;
#pragma unroll
      for (int x = 0; x < NB4; ++x) {
        const int j = 4 * x + lk;
        if (x < jmin || j >= nao) continue;
        double2* dst = S + (int64_t)(a * nao + i) * d.nso + b * nao + j;
        double2 v = part ? make_double2(0.0, 0.0) : *dst;
        v.x += scale * sr[x];
        v.y += scale * si[x];
        *dst = v;
      }
    }
  }
}

template <int NB4, int QW>
inline size_t sigma_epi4w_lds_bytes() {
  constexpr int n4 = 4 * NB4, ng = (NB4 + 3) / 4;
  return ((size_t)2 * n4 * (n4 + 1) + (size_t)(QW - 1) * ng * NB4 * 2 * 64) * sizeof(double);
}

// Sigma(tau) += sum over y < qsplit of part[y](tau), in order of y, for the 2 tc tau of one
// mirrored epilogue launch (tau = t0 + z/2, or nts - 1 - (t0 + z/2) for odd z); per_tau =
// ns * nks * nso^2 elements, slab = nts * per_tau.
__global__ void epi_part_reduce_kernel(const double2* __restrict__ part, double2* __restrict__ Sigma, int qsplit,
                                       int64_t slab, int64_t per_tau, int t0, int tc, int nts) {
  const int64_t n = (int64_t)2 * tc * per_tau;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(e / per_tau);
    const int64_t rem = e - (int64_t)z * per_tau;
    const int tz = z >> 1, tau = (z & 1) ? nts - 1 - (t0 + tz) : t0 + tz;
    const int64_t idx = (int64_t)tau * per_tau + rem;
    double2 acc = part[idx];
    for (int y = 1; y < qsplit; ++y) {
      const double2 p = part[(int64_t)y * slab + idx];
      acc.x += p.x;
      acc.y += p.y;
    }
    double2 v = Sigma[idx];
    v.x += acc.x;
    v.y += acc.y;
    Sigma[idx] = v;
  }
}

}  // namespace gwhip
