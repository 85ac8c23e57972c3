// Orbital-size contractions of the Sigma[G] vector-Jacobian product.
//
// The backward pass needs the same two shapes of nao x nao products as the forward pass, on
// operands that are adjoints, transposes or identities of the forward ones, so these two
// kernels take every operand as a strided view:
//
//   element (r, c) of operand X at   X.p + tz*X.s_t + kidx*X.s_k + Q*X.s_q + r*X.rs + c*X.cs
//   (conjugated when X.conj), kidx = k, (k + q) or (k - q) through the k-q tables.
//
//   orb_sandwich_kernel:  X_Q = L V_Q R        for every Q   (L, R shared; either may be I)
//   orb_reduce_kernel:    O  += scale sum_Q A_Q M C_Q        (M shared; may be I)
//
// Both chain two v_mfma_f64_16x16x4f64 products in registers exactly as the forward kernels
// do: a result register r of lane l holds row 16*b + (l/16) + 4r, column l%16, which is the
// K slab (l/16) + 4r of the next product's operand.  Shared small matrices sit in LDS as
// zero-padded real/imaginary planes; per-Q operands are read straight from HBM.
#pragma once

#include "gw_mfma_kernels.hpp"

namespace gwhip {

struct OrbOp {
  const double2* p;
  int64_t s_t, s_k, s_q;
  int64_t rs, cs;
  int kmap;  // 0: k, 1: k + q (kplus[k]), 2: k - q (kminus[k])
  int conj;
};

struct OrbOut {
  double2* p;
  int64_t s_t, s_k, s_q;
  int64_t rs, cs;
  int kmap;
};

struct OrbArgs {
  int nao, NQ, nk, QPW;
  const int* kplus;   // [nk] row of the k + q table for this q
  const int* kminus;  // [nk] row of the k - q table for this q
  double scale;
  int64_t kmap_st = 0;  // k +- q table row stride per blockIdx.y (0: one q for the whole launch)
};

__device__ __forceinline__ int orb_kidx(const OrbArgs& a, int kmap, int k, int tz) {
  const int64_t r = (int64_t)tz * a.kmap_st;
  return kmap == 1 ? a.kplus[r + k] : (kmap == 2 ? a.kminus[r + k] : k);
}

__device__ __forceinline__ double2 orb_load(const double2* base, const OrbOp& op, int r, int c) {
  double2 v = base[(int64_t)r * op.rs + (int64_t)c * op.cs];
  if (op.conj) v.y = -v.y;
  return v;
}

// GEMM form of the orbital products (any nao): one view at tz, for k = blockIdx.y, copied to a
// canonical row-major nao x nao matrix out[k] (conjugated when the view is).
__global__ void orb_gather_view_kernel(OrbArgs a, OrbOp op, int tz, double2* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.nao * a.nao) return;
  const int k = blockIdx.y, r = e / a.nao, c = e - (e / a.nao) * a.nao;
  const double2* base = op.p + (int64_t)tz * op.s_t + (int64_t)orb_kidx(a, op.kmap, k, tz) * op.s_k;
  out[(int64_t)k * a.nao * a.nao + e] = orb_load(base, op, r, c);
}

// O(tz, kidx(k))(r, c) += scale * T[k][r][c]; k = blockIdx.y (a permutation of the outputs).
__global__ void orb_scatter_add_kernel(OrbArgs a, OrbOut O, int tz, const double2* __restrict__ T) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.nao * a.nao) return;
  const int k = blockIdx.y, r = e / a.nao, c = e - (e / a.nao) * a.nao;
  double2* dst = O.p + (int64_t)tz * O.s_t + (int64_t)orb_kidx(a, O.kmap, k, tz) * O.s_k + (int64_t)r * O.rs +
                 (int64_t)c * O.cs;
  const double2 t = T[(int64_t)k * a.nao * a.nao + e];
  double2 v = *dst;
  v.x += a.scale * t.x;
  v.y += a.scale * t.y;
  *dst = v;
}

// Stage a shared nao x nao operand into LDS planes [16NB][16NB+1], zero outside nao.
template <int NB>
__device__ __forceinline__ void orb_stage(const OrbArgs& a, const OrbOp& op, int tz, int k, double* Xr, double* Xi) {
  constexpr int N16 = 16 * NB, LD = N16 + 1;
  const double2* base = op.p + (int64_t)tz * op.s_t + (int64_t)orb_kidx(a, op.kmap, k, tz) * op.s_k;
  for (int e = threadIdx.x; e < N16 * N16; e += blockDim.x) {
    const int r = e / N16, c = e - (e / N16) * N16;
    double2 v = make_double2(0.0, 0.0);
    if (r < a.nao && c < a.nao) v = orb_load(base, op, r, c);
    Xr[r * LD + c] = v.x;
    Xi[r * LD + c] = v.y;
  }
}

// X_Q = L V_Q R.  grid (ceil(NQ / (4 QPW)), nk, tcount); LDS 4 planes of [16NB][16NB+1] doubles.
template <int NB, bool HAS_L, bool HAS_R>
__global__ __launch_bounds__(256, NB < 3 ? 3 : 1) void orb_sandwich_kernel(OrbArgs a, OrbOp L, OrbOp V, OrbOp R, OrbOut X) {
  static_assert(HAS_L || HAS_R, "one of L, R must be present");
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  double* Lr = lds_d;
  double* Li = Lr + PL;
  double* Rr = Li + PL;
  double* Ri = Rr + PL;
  const int nao = a.nao;
  const int tz = blockIdx.z, k = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int nslab = (nao + 3) / 4;

  if (HAS_L) orb_stage<NB>(a, L, tz, k, Lr, Li);
  if (HAS_R) orb_stage<NB>(a, R, tz, k, Rr, Ri);
  __syncthreads();
  const double2* vb0 = V.p + (int64_t)tz * V.s_t + (int64_t)orb_kidx(a, V.kmap, k, tz) * V.s_k;
  double2* xb0 = X.p + (int64_t)tz * X.s_t + (int64_t)orb_kidx(a, X.kmap, k, tz) * X.s_k;
  const double sv = V.conj ? -1.0 : 1.0;

  for (int qi = 0; qi < a.QPW; ++qi) {
    const int Q = (blockIdx.x * 4 + wave) * a.QPW + qi;
    if (Q >= a.NQ) break;
    const double2* vb = vb0 + (int64_t)Q * V.s_q;
    double2* xb = xb0 + (int64_t)Q * X.s_q;
    for (int bn = 0; bn < NB; ++bn) {
      const int n = 16 * bn + lrow;
      v4d xr[NB], xi[NB];
#pragma unroll
      for (int x = 0; x < NB; ++x) { xr[x] = v4zero(); xi[x] = v4zero(); }
      if (HAS_R) {
        // W[:, bn] = V_Q R[:, bn]
        v4d wr[NB], wi[NB];
#pragma unroll
        for (int x = 0; x < NB; ++x) { wr[x] = v4zero(); wi[x] = v4zero(); }
        for (int sl = 0; sl < nslab; ++sl) {
          const int m = 4 * sl + lk;
          const double br = Rr[m * LD + n], bi = Ri[m * LD + n];
#pragma unroll
          for (int bp = 0; bp < NB; ++bp) {
            const int p = 16 * bp + lrow;
            double2 av = make_double2(0.0, 0.0);
            if (p < nao && m < nao) av = vb[(int64_t)p * V.rs + (int64_t)m * V.cs];
            cmfma(wr[bp], wi[bp], av.x, sv * av.y, br, bi);
          }
        }
        if (HAS_L) {
          // X[:, bn] = L W[:, bn]; register r of W holds row p = 16 bp + lk + 4r
#pragma unroll
          for (int bp = 0; bp < NB; ++bp)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              if (16 * bp + 4 * r >= nao) continue;  // K slab wholly in the padding
              const int p = 16 * bp + lk + 4 * r;
#pragma unroll
              for (int bt = 0; bt < NB; ++bt) {
                const int t = 16 * bt + lrow;
                cmfma(xr[bt], xi[bt], Lr[t * LD + p], Li[t * LD + p], wr[bp][r], wi[bp][r]);
              }
            }
        } else {
#pragma unroll
          for (int x = 0; x < NB; ++x) { xr[x] = wr[x]; xi[x] = wi[x]; }
        }
      } else {
        // X[:, bn] = L V_Q[:, bn]
        for (int sl = 0; sl < nslab; ++sl) {
          const int p = 4 * sl + lk;
          double2 bv = make_double2(0.0, 0.0);
          if (p < nao && n < nao) bv = vb[(int64_t)p * V.rs + (int64_t)n * V.cs];
#pragma unroll
          for (int bt = 0; bt < NB; ++bt) {
            const int t = 16 * bt + lrow;
            cmfma(xr[bt], xi[bt], Lr[t * LD + p], Li[t * LD + p], bv.x, sv * bv.y);
          }
        }
      }
      if (n < nao) {
#pragma unroll
        for (int bt = 0; bt < NB; ++bt)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int t = 16 * bt + lk + 4 * r;
            if (t < nao) xb[(int64_t)t * X.rs + (int64_t)n * X.cs] = make_double2(xr[bt][r], xi[bt][r]);
          }
      }
    }
  }
}

// O(tz, kidx(k)) += scale * sum_Q A_Q M C_Q.  grid (nk, tcount); wave w takes Q = w, w+4, ...
// LDS: M planes (2 x [16NB][16NB+1] doubles) then the cross-wave reduction buffer.
// Every (k, tz) block owns its output matrix: the k map of O must be one-to-one.
template <int NB, bool HAS_M>
__global__ __launch_bounds__(256, NB < 3 ? 3 : 1) void orb_reduce_kernel(OrbArgs a, OrbOp A, OrbOp M, OrbOp C, OrbOut O) {
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  double* Mr = lds_d;
  double* Mi = Mr + PL;
  double* red = lds_d + 2 * PL;  // [NB*NB*8][64]
  const int nao = a.nao;
  const int k = blockIdx.x, tz = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int nslab = (nao + 3) / 4;

  if (HAS_M) orb_stage<NB>(a, M, tz, k, Mr, Mi);
  __syncthreads();
  const double2* ab0 = A.p + (int64_t)tz * A.s_t + (int64_t)orb_kidx(a, A.kmap, k, tz) * A.s_k;
  const double2* cb0 = C.p + (int64_t)tz * C.s_t + (int64_t)orb_kidx(a, C.kmap, k, tz) * C.s_k;
  const double sa = A.conj ? -1.0 : 1.0, sc = C.conj ? -1.0 : 1.0;

  v4d acr[NB][NB], aci[NB][NB];
#pragma unroll
  for (int x = 0; x < NB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) { acr[x][y] = v4zero(); aci[x][y] = v4zero(); }

  for (int Q = wave; Q < a.NQ; Q += 4) {
    const double2* ab = ab0 + (int64_t)Q * A.s_q;
    const double2* cb = cb0 + (int64_t)Q * C.s_q;
    if (HAS_M) {
      // all row blocks at once (as the Sigma epilogue): each M element from LDS and each C element
      // loaded once per Q serves every row block, NB^2 independent accumulator pairs in the second
      // product; the same sums in the same order as row block by row block
      v4d tr[NB][NB], ti[NB][NB];  // [bi][bn]
#pragma unroll
      for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) { tr[x][y] = v4zero(); ti[x][y] = v4zero(); }
      for (int sl = 0; sl < nslab; ++sl) {
        const int m = 4 * sl + lk;
        double2 av[NB];
#pragma unroll
        for (int bi = 0; bi < NB; ++bi) {
          const int i = 16 * bi + lrow;
          av[bi] = (m < nao && i < nao) ? ab[(int64_t)i * A.rs + (int64_t)m * A.cs] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int bn = 0; bn < NB; ++bn) {
          const int n = 16 * bn + lrow;
          const double mr = Mr[m * LD + n], mi = Mi[m * LD + n];
#pragma unroll
          for (int bi = 0; bi < NB; ++bi) cmfma(tr[bi][bn], ti[bi][bn], mr, mi, av[bi].x, sa * av[bi].y);
        }
      }
#pragma unroll
      for (int bn = 0; bn < NB; ++bn)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (16 * bn + 4 * r >= nao) continue;  // K slab wholly in the padding
          const int n = 16 * bn + lk + 4 * r;
          double2 cv[NB];
#pragma unroll
          for (int bj = 0; bj < NB; ++bj) {
            const int j = 16 * bj + lrow;
            cv[bj] = (n < nao && j < nao) ? cb[(int64_t)n * C.rs + (int64_t)j * C.cs] : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int bi = 0; bi < NB; ++bi)
#pragma unroll
            for (int bj = 0; bj < NB; ++bj)
              cmfma(acr[bi][bj], aci[bi][bj], tr[bi][bn][r], ti[bi][bn][r], cv[bj].x, sc * cv[bj].y);
        }
      continue;
    }
#pragma unroll
    for (int bi = 0; bi < NB; ++bi) {
      const int i = 16 * bi + lrow;
      if (HAS_M) {
        // T^T = M^T A_Q^T: result register r holds T(i, n = 16 bn + lk + 4r)
        v4d tr[NB], ti[NB];
#pragma unroll
        for (int x = 0; x < NB; ++x) { tr[x] = v4zero(); ti[x] = v4zero(); }
        for (int sl = 0; sl < nslab; ++sl) {
          const int m = 4 * sl + lk;
          double2 av = make_double2(0.0, 0.0);
          if (m < nao && i < nao) av = ab[(int64_t)i * A.rs + (int64_t)m * A.cs];
#pragma unroll
          for (int bn = 0; bn < NB; ++bn) {
            const int n = 16 * bn + lrow;
            cmfma(tr[bn], ti[bn], Mr[m * LD + n], Mi[m * LD + n], av.x, sa * av.y);
          }
        }
#pragma unroll
        for (int bn = 0; bn < NB; ++bn)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (16 * bn + 4 * r >= nao) continue;  // K slab wholly in the padding
            const int n = 16 * bn + lk + 4 * r;
#pragma unroll
            for (int bj = 0; bj < NB; ++bj) {
              const int j = 16 * bj + lrow;
              double2 cv = make_double2(0.0, 0.0);
              if (n < nao && j < nao) cv = cb[(int64_t)n * C.rs + (int64_t)j * C.cs];
              cmfma(acr[bi][bj], aci[bi][bj], tr[bn][r], ti[bn][r], cv.x, sc * cv.y);
            }
          }
      } else {
        for (int sl = 0; sl < nslab; ++sl) {
          const int n = 4 * sl + lk;
          double2 av = make_double2(0.0, 0.0);
          if (i < nao && n < nao) av = ab[(int64_t)i * A.rs + (int64_t)n * A.cs];
#pragma unroll
          for (int bj = 0; bj < NB; ++bj) {
            const int j = 16 * bj + lrow;
            double2 cv = make_double2(0.0, 0.0);
            if (n < nao && j < nao) cv = cb[(int64_t)n * C.rs + (int64_t)j * C.cs];
            cmfma(acr[bi][bj], aci[bi][bj], av.x, sa * av.y, cv.x, sc * cv.y);
          }
        }
      }
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
            red[(((x * NB + y) * 2 + 0) * 4 + r) * 64 + lane] = acr[x][y][r];
            red[(((x * NB + y) * 2 + 1) * 4 + r) * 64 + lane] = aci[x][y][r];
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
            acr[x][y][r] += red[(((x * NB + y) * 2 + 0) * 4 + r) * 64 + lane];
            aci[x][y][r] += red[(((x * NB + y) * 2 + 1) * 4 + r) * 64 + lane];
          }
    }
  }
  if (wave == 0) {
    double2* ob = O.p + (int64_t)tz * O.s_t + (int64_t)orb_kidx(a, O.kmap, k, tz) * O.s_k;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y) {
        const int j = 16 * y + lrow;
        if (j >= nao) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * x + lk + 4 * r;
          if (i >= nao) continue;
          double2* dst = ob + (int64_t)i * O.rs + (int64_t)j * O.cs;
          double2 v = *dst;
          v.x += a.scale * acr[x][y][r];
          v.y += a.scale * aci[x][y][r];
          *dst = v;
        }
      }
  }
}

inline size_t orb_sandwich_lds(int NB) { return (size_t)4 * (16 * NB) * (16 * NB + 1) * sizeof(double); }
inline size_t orb_reduce_lds(int NB) {
  return ((size_t)2 * (16 * NB) * (16 * NB + 1) + (size_t)NB * NB * 8 * 64) * sizeof(double);
}

}  // namespace gwhip
