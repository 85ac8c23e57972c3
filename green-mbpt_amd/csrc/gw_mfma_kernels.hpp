// Orbital-size (nao^3 * NQ) contractions of the GW path on FP64 matrix cores.
//
// Both kernels chain two v_mfma_f64_16x16x4f64 products without an LDS round trip: for a
// 16x16 result block the MFMA leaves row (l/16) + 4r, column l%16 in register r of lane l,
// which is exactly the operand layout a following MFMA expects for its K-slab r (A operand:
// row l%16, k l/16; B operand: column l%16, k l/16).  Orbital dimensions are padded to
// NB blocks of 16 (nao <= 16 NB); complex products are four real MFMAs.
//
// Layouts (complex = double2):
//   Vt[k][Q][m][p]  = V^Q_pm(k, k + q)      per-Q transpose of the resident integrals
//   Z [tz][k][Q][t][n]                       P0 left operand (row-major per Q)
//   R [tz][k][Q][n][j]                       P-tilde conj(V), per-Q transposed
#pragma once

#include "common.hpp"
#include "gw_kernels.hpp"

namespace gwhip {

// (cr, ci) += (ar + i ai) * (br + i bi) on 16x16x4 fragments
__device__ __forceinline__ void cmfma(v4d& cr, v4d& ci, double ar, double ai, double br, double bi) {
  cr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, ci, 0, 0, 0);
}

__device__ __forceinline__ v4d v4zero() { return v4d{0.0, 0.0, 0.0, 0.0}; }

// ---------------------------------------------------------------------------------------
// P0 left operand  Z_Q = sum_s Gb_s V_Q Gq_s  (Gb_s = G(beta - tau, k) block (b, a),
// Gq_s = G(tau, k + q) block (a, b)); src/gw_cpu_kernel.cpp:170-213 regrouped.
// grid: (ceil(NQ / (4 QPW)), nk, tcount); 4 waves, each handles QPW consecutive Q.
// LDS per spin block: Gb and Gq as zero-padded real/imag planes [16NB][16NB+1].
template <int NB>
__global__ __launch_bounds__(256) void z_prep_mfma_kernel(DevDims d, const double2* __restrict__ G,
                                                          const double2* __restrict__ Vt,
                                                          double2* __restrict__ Z, const int* __restrict__ kplus,
                                                          int t0, int QPW) {
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  double* Gbr = lds_d;
  double* Gbi = Gbr + PL;
  double* Gqr = Gbi + PL;
  double* Gqi = Gqr + PL;
  const int nao = d.nao, nao2 = nao * nao;
  const int tz = blockIdx.z, tau = t0 + tz, tt = d.nts - 1 - tau;
  const int k = blockIdx.y, kq = kplus[k];
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
      Gbr[r * LD + c] = gb.x;
      Gbi[r * LD + c] = gb.y;
      Gqr[r * LD + c] = gq.x;
      Gqi[r * LD + c] = gq.y;
    }
    __syncthreads();
    for (int qi = 0; qi < QPW; ++qi) {
      const int Q = (blockIdx.x * 4 + wave) * QPW + qi;
      if (Q >= d.NQ) break;
      const double2* vt = Vt + ((int64_t)k * d.NQ + Q) * nao2;  // vt[m * nao + p] = V[p][m]
      double2* zq = Z + (((int64_t)tz * d.nk + k) * d.NQ + Q) * nao2;
      for (int bn = 0; bn < NB; ++bn) {
        // W[:, bn] = V Gq[:, bn]
        v4d wr[NB], wi[NB];
#pragma unroll
        for (int x = 0; x < NB; ++x) { wr[x] = v4zero(); wi[x] = v4zero(); }
        for (int sl = 0; sl < nslab; ++sl) {
          const int m = 4 * sl + lk;
          const double br = Gqr[m * LD + 16 * bn + lrow], bi = Gqi[m * LD + 16 * bn + lrow];
#pragma unroll
          for (int bp = 0; bp < NB; ++bp) {
            const int p = 16 * bp + lrow;
            double2 av = make_double2(0.0, 0.0);
            if (p < nao && m < nao) av = vt[m * nao + p];
            cmfma(wr[bp], wi[bp], av.x, av.y, br, bi);
          }
        }
        // Z[:, bn] (+)= Gb W[:, bn]
        v4d zr[NB], zi[NB];
#pragma unroll
        for (int x = 0; x < NB; ++x) { zr[x] = v4zero(); zi[x] = v4zero(); }
#pragma unroll
        for (int bp = 0; bp < NB; ++bp)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int p = 16 * bp + 4 * r + lk;
            const double bwr = wr[bp][r], bwi = wi[bp][r];
#pragma unroll
            for (int bt = 0; bt < NB; ++bt) {
              const int t = 16 * bt + lrow;
              cmfma(zr[bt], zi[bt], Gbr[t * LD + p], Gbi[t * LD + p], bwr, bwi);
            }
          }
        const int n = 16 * bn + lrow;
        if (n < nao) {
#pragma unroll
          for (int bt = 0; bt < NB; ++bt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int t = 16 * bt + lk + 4 * r;
              if (t < nao) {
                double2 v = make_double2(zr[bt][r], zi[bt][r]);
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
}

// ---------------------------------------------------------------------------------------
// Sigma(tau, is, k)[a,b](i, j) += scale * sum_Q sum_n T_Q[i][n] R_Q[n][j],  T_Q = V_Q G_s
// with V = V(k, k - q), G_s = G(tau, k - q) block (a, b)  (src/gw_cpu_kernel.cpp:359-373,
// Y2 = Y1^T P done by the big GEMM as R = P-tilde conj(V)).  The first MFMA produces
// T^T = G^T V^T so that its result registers feed the second MFMA's A operand.
// grid: (nk, tcount); wave w takes Q = w, w + 4, ...; the four partial sums meet in LDS.
template <int NB>
__global__ __launch_bounds__(256) void sigma_epi_mfma_kernel(DevDims d, const double2* __restrict__ G,
                                                             const double2* __restrict__ Vt,
                                                             const double2* __restrict__ R,
                                                             double2* __restrict__ Sigma,
                                                             const int* __restrict__ kminus, int t0, double scale) {
  constexpr int N16 = 16 * NB, LD = N16 + 1, PL = N16 * LD;
  extern __shared__ double lds_d[];
  double* Gr = lds_d;
  double* Gi = Gr + PL;
  double* red = Gi + PL;  // [NB*NB*8][64]
  const int nao = d.nao, nao2 = nao * nao;
  const int k = blockIdx.x, tz = blockIdx.y, tau = t0 + tz;
  const int kmq = kminus[k];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  const int nslab = (nao + 3) / 4;

  for (int s = 0; s < d.pseudo_ns; ++s) {
    const int is = s / (d.pns * d.pns), a = (s / d.pns) % d.pns, b = s % d.pns;
    const double2* Gk = G + g_index(d, tau, is, kmq);
    __syncthreads();
    for (int e = tid; e < N16 * N16; e += blockDim.x) {
      const int r = e / N16, c = e - (e / N16) * N16;
      double2 g = make_double2(0.0, 0.0);
      if (r < nao && c < nao) g = Gk[(int64_t)(a * nao + r) * d.nso + b * nao + c];  // G[m = r][n = c]
      Gr[r * LD + c] = g.x;
      Gi[r * LD + c] = g.y;
    }
    __syncthreads();
    v4d sr[NB][NB], si[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y) { sr[x][y] = v4zero(); si[x][y] = v4zero(); }

    for (int Q = wave; Q < d.NQ; Q += 4) {
      const double2* vt = Vt + ((int64_t)k * d.NQ + Q) * nao2;                   // vt[m * nao + i] = V[i][m]
      const double2* rq = R + (((int64_t)tz * d.nk + k) * d.NQ + Q) * nao2;     // rq[n * nao + j]
#pragma unroll
      for (int bi = 0; bi < NB; ++bi) {
        v4d tr[NB], ti[NB];
#pragma unroll
        for (int x = 0; x < NB; ++x) { tr[x] = v4zero(); ti[x] = v4zero(); }
        const int i = 16 * bi + lrow;
        for (int sl = 0; sl < nslab; ++sl) {
          const int m = 4 * sl + lk;
          double2 bv = make_double2(0.0, 0.0);
          if (m < nao && i < nao) bv = vt[m * nao + i];
#pragma unroll
          for (int bn = 0; bn < NB; ++bn) {
            const int n = 16 * bn + lrow;
            cmfma(tr[bn], ti[bn], Gr[m * LD + n], Gi[m * LD + n], bv.x, bv.y);
          }
        }
#pragma unroll
        for (int bn = 0; bn < NB; ++bn)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int n = 16 * bn + 4 * r + lk;
            const double ar = tr[bn][r], ai = ti[bn][r];
#pragma unroll
            for (int bj = 0; bj < NB; ++bj) {
              const int j = 16 * bj + lrow;
              double2 rv = make_double2(0.0, 0.0);
              if (n < nao && j < nao) rv = rq[n * nao + j];
              cmfma(sr[bi][bj], si[bi][bj], ar, ai, rv.x, rv.y);
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
      double2* S = Sigma + g_index(d, tau, is, k);
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
            double2* dst = S + (int64_t)(a * nao + i) * d.nso + b * nao + j;
            double2 v = *dst;
            v.x += scale * sr[x][y][r];
            v.y += scale * si[x][y][r];
            *dst = v;
          }
        }
    }
  }
}

// Per-Q transpose of the resident integrals: Vt[..][Q][b][a] = V[..][Q][a][b].
__global__ void transpose_v_kernel(const double2* __restrict__ V, double2* __restrict__ Vt, int nao, int64_t nmat) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nao2 = (int64_t)nao * nao;
  if (gid >= nmat * nao2) return;
  const int64_t mat = gid / nao2;
  const int e = (int)(gid - mat * nao2);
  const int a = e / nao, b = e - (e / nao) * nao;
  Vt[mat * nao2 + (int64_t)b * nao + a] = V[gid];
}

}  // namespace gwhip
