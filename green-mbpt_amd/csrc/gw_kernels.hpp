// GW-specific device kernels (FP64).  Index conventions (all row-major, complex = double2):
//   G, Sigma : [nts][ns][nk][nso][nso]
//   Vq       : [nq][nk][NQ][nao][nao]   Vq[q][k] = V^Q(k, k+q)   (every ordered k pair resident)
//   P        : [nq][nts][NQ][NQ]         P0 on the first nts/2 tau, then P-tilde on all nts
//   Vt       : as Vq, each Q block transposed (Vt[q][k][Q][b][a] = Vq[q][k][Q][a][b])
//   Z        : [tc][nk][NQ][nao][nao]    Z = sum_s Gbar_s V G_s   (P0 left operand)
//   R        : [tc][nk][NQ][nao][nao]    R = P-tilde conj(V), each Q block transposed
// Spin blocks follow the reference (src/gw_cpu_kernel.cpp:166-186, 328-351): block s of
// pseudo_ns has is = s/(pns^2), a = (s/pns)%pns, b = s%pns.
#pragma once

#include "common.hpp"

namespace gwhip {

struct DevDims {
  int nts, ns, nk, nso, nao, NQ, pns, pseudo_ns;
};

__device__ __forceinline__ int64_t g_index(const DevDims& d, int t, int is, int k) {
  return (((int64_t)t * d.ns + is) * d.nk + k) * d.nso * d.nso;
}

// ---------------------------------------------------------------------------------------
// X = (X + X^H)/2 for a batch of N x N matrices (common_defs.h:104-115), with coalesced accesses:
// block (y = matrix) takes the tile pair (I, J), I <= J,
// of 32 x 32 tiles (blockIdx.x enumerates the upper tile triangle), stages tile (I, J) and tile (J, I)
// in LDS, and writes h = (A[a][b] + conj(B[b][a]))/2 to (I, J) and conj(h) to (J, I) (on a diagonal
// tile the same formula is Hermitian by construction, with a real diagonal).
constexpr int HERM_TILE = 32;
__global__ __launch_bounds__(256) void hermitize_tiled_kernel(double2* X, int N, int64_t stride) {
  __shared__ double2 ta[HERM_TILE][HERM_TILE + 1], tb[HERM_TILE][HERM_TILE + 1];
  const int T = (N + HERM_TILE - 1) / HERM_TILE;
  // blockIdx.x -> (I, J) with I <= J, row-major over the upper triangle
  int I = 0, rem = (int)blockIdx.x;
  while (rem >= T - I) {
    rem -= T - I;
    ++I;
  }
  const int J = I + rem;
  double2* M = X + (int64_t)blockIdx.y * stride;
  const int i0 = I * HERM_TILE, j0 = J * HERM_TILE;
  for (int e = threadIdx.x; e < HERM_TILE * HERM_TILE; e += blockDim.x) {
    const int a = e / HERM_TILE, b = e % HERM_TILE;
    if (i0 + a < N && j0 + b < N) ta[a][b] = M[(int64_t)(i0 + a) * N + j0 + b];
    if (j0 + a < N && i0 + b < N) tb[a][b] = M[(int64_t)(j0 + a) * N + i0 + b];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < HERM_TILE * HERM_TILE; e += blockDim.x) {
    const int a = e / HERM_TILE, b = e % HERM_TILE;
    if (i0 + a >= N || j0 + b >= N) continue;
    const double2 u = ta[a][b], l = tb[b][a];
    ta[a][b] = make_double2(0.5 * (u.x + l.x), 0.5 * (u.y - l.y));
  }
  __syncthreads();
  for (int e = threadIdx.x; e < HERM_TILE * HERM_TILE; e += blockDim.x) {
    const int a = e / HERM_TILE, b = e % HERM_TILE;
    if (i0 + a < N && j0 + b < N) M[(int64_t)(i0 + a) * N + j0 + b] = ta[a][b];
    if (I != J && j0 + a < N && i0 + b < N) {
      const double2 h = ta[b][a];
      M[(int64_t)(j0 + a) * N + i0 + b] = make_double2(h.x, -h.y);
    }
  }
}

// (X + X^H)/2 for nmat N x N matrices at `stride` apart (nmat <= 65535).
inline void hermitize(double2* X, int N, int64_t stride, int64_t nmat, hipStream_t st) {
  const int T = (N + HERM_TILE - 1) / HERM_TILE;
  hipLaunchKernelGGL(hermitize_tiled_kernel, dim3((unsigned)(T * (T + 1) / 2), (unsigned)nmat), dim3(256), 0, st, X, N,
                     stride);
}

// Upper triangle -> Hermitian matrix: X[i][j] = conj(X[j][i]) below the diagonal, real diagonal.
// Used when only the upper tiles of a Hermitian product were computed (the result equals
// (X + X^H)/2 up to rounding when X is Hermitian).  One work-group per 32 x 32 tile (I, J), I <= J,
// of the upper triangle, staged in LDS so that both the read of (I, J) and the write of (J, I) are
// row-contiguous.
__global__ __launch_bounds__(256) void mirror_upper_tiled_kernel(double2* X, int N, int64_t stride) {
  __shared__ double2 ta[HERM_TILE][HERM_TILE + 1];
  const int T = (N + HERM_TILE - 1) / HERM_TILE;
  int I = 0, rem = (int)blockIdx.x;
  while (rem >= T - I) {
    rem -= T - I;
    ++I;
  }
  const int J = I + rem;
  double2* M = X + (int64_t)blockIdx.y * stride;
  const int i0 = I * HERM_TILE, j0 = J * HERM_TILE;
  for (int e = threadIdx.x; e < HERM_TILE * HERM_TILE; e += blockDim.x) {
    const int a = e / HERM_TILE, b = e % HERM_TILE;
    if (i0 + a < N && j0 + b < N) ta[a][b] = M[(int64_t)(i0 + a) * N + j0 + b];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < HERM_TILE * HERM_TILE; e += blockDim.x) {
    const int a = e / HERM_TILE, b = e % HERM_TILE;  // element (j0 + a, i0 + b) of the lower tile
    if (j0 + a >= N || i0 + b >= N) continue;
    if (I == J && b > a) continue;  // diagonal tile: its upper part stays
    const double2 u = ta[b][a];
    M[(int64_t)(j0 + a) * N + i0 + b] = make_double2(u.x, a == b && I == J ? 0.0 : -u.y);
  }
}

// Mirror the upper triangles of nmat N x N matrices at `stride` apart (nmat <= 65535).
inline void mirror_upper(double2* X, int N, int64_t stride, int64_t nmat, hipStream_t st) {
  const int T = (N + HERM_TILE - 1) / HERM_TILE;
  hipLaunchKernelGGL(mirror_upper_tiled_kernel, dim3((unsigned)(T * (T + 1) / 2), (unsigned)nmat), dim3(256), 0, st, X,
                     N, stride);
}

// max_(i,j) |X[i][j] - conj(X[j][i])| and max |X| over nmat N x N matrices, as the bit patterns
// of non-negative doubles (ordered like the values) combined with atomicMax.
__global__ void herm_check_kernel(const double2* X, int64_t nmat, int N, unsigned long long* out) {
  const int64_t n2 = (int64_t)N * N, total = nmat * n2;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  double dmax = 0.0, amax = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += step) {
    const int64_t m = e / n2, r = e - m * n2;
    const int i = (int)(r / N), j = (int)(r - (r / N) * N);
    const double2 a = X[e], b = X[m * n2 + (int64_t)j * N + i];
    dmax = fmax(dmax, fmax(fabs(a.x - b.x), fabs(a.y + b.y)));
    amax = fmax(amax, fmax(fabs(a.x), fabs(a.y)));
  }
  for (int o = 32; o > 0; o >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
    amax = fmax(amax, __shfl_xor(amax, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out + 0, (unsigned long long)__double_as_longlong(dmax));
    atomicMax(out + 1, (unsigned long long)__double_as_longlong(amax));
  }
}

// P0 half-grid symmetrisation for all q: Hermitise P[q][t] for t < nts/2
// (src/gw_cpu_kernel.cpp:217-225; the /nk and the prefactor are applied by the GEMM).
// The mirror P0(beta - tau) = P0(tau) is folded into the tau -> Omega matrix instead.

// Dyson operand: X[b] = I - Pw[b]
__global__ void identity_minus_kernel(double2* X, const double2* Pw, int N, int64_t batch) {
  const int64_t n2 = (int64_t)N * N;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n2 * batch) return;
  const int64_t e = gid % n2;
  const int i = (int)(e / N), j = (int)(e - (e / N) * N);
  const double2 p = Pw[gid];
  X[gid] = make_double2((i == j ? 1.0 : 0.0) - p.x, -p.y);
}

// Hermitian block Gauss-Jordan sweep (the sweep operator on a Hermitian matrix).  After the blocks
// S = [0, J0) are swept, A holds
//     [ A_SS^-1            A_SS^-1 A_SU            ]     X = A_SS^-1 A_SU,
//     [ -(X)^H             A_UU - A_US A_SS^-1 A_SU ]
// i.e. Hermitian diagonal parts and anti-Hermitian-paired off-diagonal parts, so the upper triangle
// determines the matrix and the sweep keeps only the upper (block) triangle up to date; after the
// last block it holds the upper triangle of A^-1.
// Step J (columns [J0, J0 + jb)), data movement only: the panel
//   P[m][l][i] = A[J0 + l][i]          for i >= J0  (pivot rows, upper triangle)
//              = conj(A[i][J0 + l])    for i <  J0  (swept rows' column J, upper triangle)
// and the pivot block D[m][l][l'] = A[J0 + l][J0 + l'] (a diagonal tile, both triangles current).
// Then the old column panel is P^H and the old pivot rows are P with the columns i < J0 negated.
__global__ void hgj_panel_kernel(const double2* __restrict__ A, int N, int64_t sA, int J0, int jb,
                                 double2* __restrict__ P, int64_t sP, double2* __restrict__ D, int64_t sD) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)N * jb) return;
  const int l = (int)(e / N), i = (int)(e - (e / N) * N);
  const int64_t m = blockIdx.y;
  double2 v;
  if (i >= J0) {
    v = A[m * sA + (int64_t)(J0 + l) * N + i];
  } else {
    const double2 u = A[m * sA + (int64_t)i * N + J0 + l];
    v = make_double2(u.x, -u.y);
  }
  P[m * sP + (int64_t)l * N + i] = v;
  if (i >= J0 && i < J0 + jb) D[m * sD + (int64_t)l * jb + (i - J0)] = v;
}

// Q[m][l][i] = -Q[m][l][i] for i < J0: D^-1 P becomes the new pivot rows D^-1 A_J* (the old pivot
// rows at swept columns are -conj of the stored column panel).
__global__ void hgj_negate_kernel(double2* __restrict__ Q, int N, int64_t sQ, int J0, int jb) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)jb * J0) return;
  const int l = (int)(e / J0), i = (int)(e - (e / J0) * J0);
  double2* q = Q + blockIdx.y * sQ + (int64_t)l * N + i;
  *q = make_double2(-q->x, -q->y);
}

// The pivot rows of step J in the upper triangle: A[J0 + l][j] = Q[l][j] for j >= J0 + jb and
// A[J0 + l][J0 + l'] = D[l][l'] (the inverted pivot block).
__global__ void hgj_rows_kernel(double2* __restrict__ A, int N, int64_t sA, int J0, int jb,
                                const double2* __restrict__ Q, int64_t sQ, const double2* __restrict__ D, int64_t sD) {
  const int W = N - J0;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)jb * W) return;
  const int l = (int)(e / W), j = J0 + (int)(e - (e / W) * W);
  const int64_t m = blockIdx.y;
  A[m * sA + (int64_t)(J0 + l) * N + j] = j < J0 + jb ? D[m * sD + (int64_t)l * jb + (j - J0)] : Q[m * sQ + (int64_t)l * N + j];
}

// ---------------------------------------------------------------------------------------
// Batched in-place inverse of Hermitian positive-definite N x N matrices by blocked
// Gauss-Jordan sweeps without pivoting (the reference uses Eigen's LDLT on I - P0,
// src/gw_cpu_kernel.cpp:243-252; for the polarisation of a physical G, I - P0(i Omega) is
// positive definite, so every pivot is positive).  One workgroup per matrix; per block column J:
//   Dinv = A_JJ^-1 (small sweep in LDS);  R = Dinv A_JI;  C = A_IJ (old);
//   A_II -= C R;  A_IJ = -C Dinv;  A_JI = R;  A_JJ = Dinv.
// Threads own columns; the column panel C lives in LDS ([N][nb]) and the row panel in registers.
// Without pivoting the k-th pivot is the k-th leading Schur complement's corner, real and positive
// exactly when the matrix is Hermitian positive definite; a pivot whose real part is not positive
// (or not finite) sets *status to 1, so that the caller reports the matrix instead of returning a
// silently degraded inverse.  The Dyson step (forward and VJP) runs it on the 64 x 64 pivot blocks
// of the Hermitian block sweep above.
constexpr int GJ_NB = 16;

__global__ void gj_inverse_kernel(double2* A, int N, int64_t stride, int nb, int* status) {
  extern __shared__ double2 lds[];
  double2* D = lds;               // [GJ_NB][GJ_NB]
  double2* Cp = D + GJ_NB * GJ_NB;  // [N][nb]
  double2* M = A + (int64_t)blockIdx.x * stride;
  const int tid = threadIdx.x, nth = blockDim.x;

  for (int J0 = 0; J0 < N; J0 += nb) {
    const int jb = min(nb, N - J0);
    __syncthreads();
    for (int e = tid; e < jb * jb; e += nth) {
      const int r = e / jb, c = e - (e / jb) * jb;
      D[r * GJ_NB + c] = M[(int64_t)(J0 + r) * N + J0 + c];
    }
    for (int e = tid; e < N * jb; e += nth) {
      const int i = e / jb, l = e - (e / jb) * jb;
      Cp[i * nb + l] = M[(int64_t)i * N + J0 + l];
    }
    __syncthreads();
    // small in-place sweep of D (jb x jb); the launcher guarantees nth >= GJ_NB*GJ_NB so
    // each element has its own thread: read old values, barrier, write.
    for (int piv = 0; piv < jb; ++piv) {
      const bool act = tid < jb * jb;
      double2 v = make_double2(0.0, 0.0);
      int r = 0, c = 0;
      if (act) {
        r = tid / jb;
        c = tid - r * jb;
        const double2 pv = D[piv * GJ_NB + piv];
        if (tid == 0 && !(pv.x > 0.0 && fabs(pv.y) <= pv.x)) atomicOr(status, 1);
        const double den = pv.x * pv.x + pv.y * pv.y;
        const double2 pinv = make_double2(pv.x / den, -pv.y / den);
        if (r == piv && c == piv) {
          v = pinv;
        } else if (r == piv) {
          v = cmul(D[piv * GJ_NB + c], pinv);
        } else if (c == piv) {
          v = cmul(D[r * GJ_NB + piv], pinv);
          v.x = -v.x;
          v.y = -v.y;
        } else {
          const double2 t = cmul(cmul(D[r * GJ_NB + piv], D[piv * GJ_NB + c]), pinv);
          const double2 o = D[r * GJ_NB + c];
          v = make_double2(o.x - t.x, o.y - t.y);
        }
      }
      __syncthreads();
      if (act) D[r * GJ_NB + c] = v;
      __syncthreads();
    }
    // column sweep
    for (int j = tid; j < N; j += nth) {
      double2 rv[GJ_NB];
      const bool inJ = (j >= J0) && (j < J0 + jb);
      if (inJ) {
#pragma unroll
        for (int l = 0; l < GJ_NB; ++l) rv[l] = (l < jb) ? D[l * GJ_NB + (j - J0)] : make_double2(0.0, 0.0);
      } else {
        double2 col[GJ_NB];
#pragma unroll
        for (int l = 0; l < GJ_NB; ++l) col[l] = (l < jb) ? M[(int64_t)(J0 + l) * N + j] : make_double2(0.0, 0.0);
#pragma unroll
        for (int l = 0; l < GJ_NB; ++l) {
          double2 acc = make_double2(0.0, 0.0);
          if (l < jb)
            for (int lp = 0; lp < jb; ++lp) acc = cfma(D[l * GJ_NB + lp], col[lp], acc);
          rv[l] = acc;
        }
      }
#pragma unroll
      for (int l = 0; l < GJ_NB; ++l)
        if (l < jb) M[(int64_t)(J0 + l) * N + j] = rv[l];
      for (int i = 0; i < N; ++i) {
        if (i >= J0 && i < J0 + jb) continue;
        double2 v = inJ ? make_double2(0.0, 0.0) : M[(int64_t)i * N + j];
        const double2* c = Cp + i * nb;
#pragma unroll
        for (int l = 0; l < GJ_NB; ++l)
          if (l < jb) {
            const double2 t = cmul(c[l], rv[l]);
            v.x -= t.x;
            v.y -= t.y;
          }
        M[(int64_t)i * N + j] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// V expansion: Vq[q][k][Q][i][j] = V^Q_ij(k, k+q) from the packed pair store
// (df_integral_t.h:68,169-176: k1 >= k2 stored, k1 < k2 = per-Q adjoint).
__global__ void expand_v_kernel(const double2* __restrict__ Vp, double2* __restrict__ Vq,
                                const int* __restrict__ kplus_tab,  // [nk][nq]: k + q
                                int nk, int NQ, int nao, int64_t total) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const int nao2 = nao * nao;
  int64_t r = gid;
  const int j = (int)(r % nao); r /= nao;
  const int i = (int)(r % nao); r /= nao;
  const int Q = (int)(r % NQ); r /= NQ;
  const int k = (int)(r % nk); r /= nk;
  const int q = (int)r;
  const int k2 = kplus_tab[k * nk + q];
  double2 v;
  if (k >= k2) {
    const int64_t p = (int64_t)k * (k + 1) / 2 + k2;
    v = Vp[(p * NQ + Q) * nao2 + i * nao + j];
  } else {
    const int64_t p = (int64_t)k2 * (k2 + 1) / 2 + k;
    v = cconj(Vp[(p * NQ + Q) * nao2 + j * nao + i]);
  }
  Vq[gid] = v;
}

// Counter-based normal variates (SplitMix64 + Box-Muller), identical counters to
// gwhip.inputs.normal_pairs.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t key, uint64_t idx) {
  const uint64_t u = splitmix64(key + idx);
  return ((double)(u >> 11) + 0.5) * 0x1.0p-53;
}
__device__ __forceinline__ double2 normal_pair(uint64_t key, uint64_t idx, double sigma) {
  const double u1 = u01(key, 2 * idx), u2 = u01(key, 2 * idx + 1);
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  return make_double2(sigma * r * c, sigma * r * s);
}

__global__ void generate_v_kernel(double2* __restrict__ Vq, const int* __restrict__ kplus_tab, int nk, int NQ,
                                  int nao, int64_t total, uint64_t key, double sigma) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  int64_t r = gid;
  const int j = (int)(r % nao); r /= nao;
  const int i = (int)(r % nao); r /= nao;
  const int Q = (int)(r % NQ); r /= NQ;
  const int k = (int)(r % nk); r /= nk;
  const int q = (int)r;
  const int k2 = kplus_tab[k * nk + q];
  auto raw = [&](int64_t p, int ii, int jj) {
    return normal_pair(key, ((p * NQ + Q) * nao + ii) * nao + jj, sigma);
  };
  double2 v;
  if (k == k2) {
    const int64_t p = (int64_t)k * (k + 1) / 2 + k;
    const double2 w1 = raw(p, i, j), w2 = raw(p, j, i);
    v = make_double2(0.5 * (w1.x + w2.x), 0.5 * (w1.y - w2.y));
  } else if (k > k2) {
    v = raw((int64_t)k * (k + 1) / 2 + k2, i, j);
  } else {
    v = cconj(raw((int64_t)k2 * (k2 + 1) / 2 + k, j, i));
  }
  Vq[gid] = v;
}

// Non-finite scan: flag[0] = 1 if any component is NaN/Inf.
// Fill kernels (used instead of memset so that captured graphs hold kernel nodes only).
__global__ void zero_kernel(double2* X, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) X[i] = make_double2(0.0, 0.0);
}

// Out[k] = G[base + map[k] nso^2 ...] for k < grid.y: one (tau, spin) slice of G at mapped k
// points, whole nso x nso matrices.
__global__ void gather_g_kernel(const double2* __restrict__ G, double2* __restrict__ out,
                                const int* __restrict__ map, int64_t nso2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nso2) return;
  const int64_t k = blockIdx.y;
  out[k * nso2 + e] = G[(int64_t)map[k] * nso2 + e];
}

__global__ void zero_int_kernel(int* X, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) X[i] = 0;
}

__global__ void nonfinite_kernel(const double2* X, int64_t n, int* flag) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (; gid < n; gid += step) {
    const double2 v = X[gid];
    bad |= !isfinite(v.x) || !isfinite(v.y);
  }
  if (bad) atomicOr(flag, 1);
}

// q -> 0 extrapolation (aux_to_PW_00, src/mbpt_q0_utils_t.cpp:11-20): eps[n][iq] = conj(A_iq)^T Y(n) A_iq
// for the q of this Dyson chunk that are irreducible; one block per (n, q in the chunk).
__global__ __launch_bounds__(256) void q0_eps_kernel(const double2* __restrict__ Y, const double2* __restrict__ A,
                                                     const int* __restrict__ qiq, int q0, int nwb, int NQ, int inq,
                                                     double2* __restrict__ eps) {
  const int n = blockIdx.x, qc = blockIdx.y;
  const int iq = qiq[q0 + qc];
  if (iq < 0) return;
  const int64_t NQ2 = (int64_t)NQ * NQ;
  const double2* P = Y + ((int64_t)qc * nwb + n) * NQ2;
  const double2* a = A + (int64_t)iq * NQ;
  double sr = 0.0, si = 0.0;
  for (int64_t e = threadIdx.x; e < NQ2; e += blockDim.x) {
    const int p = (int)(e / NQ), r = (int)(e - (e / NQ) * NQ);
    const double2 x = P[e], ap = a[p], ar = a[r];
    // conj(ap) * x * ar
    const double tr = ap.x * x.x + ap.y * x.y, ti = ap.x * x.y - ap.y * x.x;
    sr += tr * ar.x - ti * ar.y;
    si += tr * ar.y + ti * ar.x;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) eps[(int64_t)n * inq + iq] = make_double2(red[0][0], red[1][0]);
}

}  // namespace gwhip
