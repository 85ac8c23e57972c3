// Register-only throughput probes and a GEMM timing harness for one MI355X (gfx950): a debug
// library (libgwhip_debug.so, include/gwhip_debug.h) kept apart from libgwhip.so.  Nothing in the
// GW path calls it; it serves ceiling measurements only.
#include <hip/hip_runtime.h>

#include "../../include/gwhip.h"
#include "../../include/gwhip_debug.h"
#include "common.hpp"
#include "gw_kernels.hpp"
#include "zgemm_mfma.hpp"

namespace gwhip {
__global__ void fill_hash_kernel(double2* X, int64_t n, uint64_t salt) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  X[gid] = make_double2(u01(salt, 2 * gid) - 0.5, u01(salt, 2 * gid + 1) - 0.5);
}

// v_mfma_f64_4x4x4_4b: four independent 4x4x4 products per instruction (64 results, one per lane)
template <int CH>
__global__ __launch_bounds__(256) void mfma4_peak_kernel(double* out, int iters) {
  double acc[CH];
#pragma unroll
  for (int j = 0; j < CH; ++j) acc[j] = 0.0;
  const double a = 1e-3 * threadIdx.x, b = 1.0 + 1e-6 * blockIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < CH; ++j) acc[j] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < CH; ++j) s += acc[j];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Lane layout of v_mfma_f64_4x4x4_4b, one wave, out[run * 64 + lane] = D of that lane:
//   run 0: A = 2^(l % 16), B = 2^(16 (l % 16))  -> each D is a sum of 2^(a + 16 b) over its k pairs
//   run 1: A = l / 16 + 1, B = 1                 -> 4 (block of A feeding lane l) when blocks are l / 16
//   run 2: A = 1, B = l / 16 + 1
__global__ __launch_bounds__(64) void mfma4_layout_kernel(double* out) {
  const int l = threadIdx.x;
  const double a0 = ldexp(1.0, l % 16), b0 = ldexp(1.0, 16 * (l % 16));
  out[l] = __builtin_amdgcn_mfma_f64_4x4x4f64(a0, b0, 0.0, 0, 0, 0);
  out[64 + l] = __builtin_amdgcn_mfma_f64_4x4x4f64((double)(l / 16 + 1), 1.0, 0.0, 0, 0, 0);
  out[128 + l] = __builtin_amdgcn_mfma_f64_4x4x4f64(1.0, (double)(l / 16 + 1), 0.0, 0, 0, 0);
}

template <int CH>
__global__ __launch_bounds__(256) void mfma_peak_kernel(double* out, int iters) {
  v4d acc[CH];
#pragma unroll
  for (int j = 0; j < CH; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  const double a = 1e-3 * threadIdx.x, b = 1.0 + 1e-6 * blockIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < CH; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < CH; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64 vector FMA probe (mode 1) and MFMA + vector FMA in the same wave (mode 2).
template <int MODE>
__global__ __launch_bounds__(256) void valu_peak_kernel(double* out, int iters) {
  double x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = 1e-3 * (threadIdx.x + j);
  v4d acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  const double a = 1.0 - 1e-9 * blockIdx.x, b = 1e-7;
  for (int i = 0; i < iters; ++i) {
    if (MODE == 2) {
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = fma(x[j], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += x[j];
#pragma unroll
  for (int j = 0; j < 4; ++j) s += acc[j][0];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64 vector-FMA complex outer products with operands from LDS (probe): each lane accumulates a
// 4 x 4 complex tile over K = 32, A and B as real/imaginary planes [k][32] read 4 at a time.
__global__ __launch_bounds__(256) void valu_cgemm_probe_kernel(double* out, int iters) {
  __shared__ double Ar[32][32], Ai[32][32], Br[32][32], Bi[32][32];
  for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
    const int r = e / 32, c = e % 32;
    Ar[r][c] = 1e-3 * (r + c);
    Ai[r][c] = 1e-3 * (r - c);
    Br[r][c] = 1e-3 * (2 * r + c);
    Bi[r][c] = 1e-3 * (r - 2 * c);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, ti = (lane & 7) * 4, tj = (lane >> 3) * 4;
  double cr[4][4], ci[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) { cr[x][y] = 0.0; ci[x][y] = 0.0; }
  for (int it = 0; it < iters; ++it) {
#pragma unroll 4
    for (int k = 0; k < 32; ++k) {
      double ar[4], ai[4], br[4], bi[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        ar[x] = Ar[k][ti + x];
        ai[x] = Ai[k][ti + x];
        br[x] = Br[k][tj + x];
        bi[x] = Bi[k][tj + x];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          cr[x][y] = fma(ar[x], br[y], cr[x][y]);
          cr[x][y] = fma(-ai[x], bi[y], cr[x][y]);
          ci[x][y] = fma(ar[x], bi[y], ci[x][y]);
          ci[x][y] = fma(ai[x], br[y], ci[x][y]);
        }
    }
  }
  double s = 0.0;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) s += cr[x][y] + ci[x][y];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf_t __attribute__((ext_vector_type(8)));

// Low-precision MFMA probes: MODE 0 = v_mfma_i32_16x16x64_i8, MODE 1 = v_mfma_f32_16x16x32_bf16;
// eight independent accumulator chains per wave.
template <int MODE>
__global__ __launch_bounds__(256) void lowp_peak_kernel(double* out, int iters) {
  v4i_t ai = {(int)threadIdx.x, 3, 5, 7}, bi = {11, (int)blockIdx.x, 13, 17};
  v8bf_t ab, bb;
#pragma unroll
  for (int j = 0; j < 8; ++j) { ab[j] = (__bf16)(0.001f * (threadIdx.x + j)); bb[j] = (__bf16)(0.5f); }
  v4i_t ci[8];
  v4f_t cf[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) { ci[j] = v4i_t{0, 0, 0, 0}; cf[j] = v4f_t{0.f, 0.f, 0.f, 0.f}; }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (MODE == 0) ci[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai, bi, ci[j], 0, 0, 0);
      else cf[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab, bb, cf[j], 0, 0, 0);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += (double)ci[j][0] + (double)cf[j][0];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef int v16i_t __attribute__((ext_vector_type(16)));

// int8 MFMA shape probe: v_mfma_i32_32x32x32_i8 (65536 ops) with NCH independent 32 x 32 accumulator
// chains per wave, or (WIDE = 0) v_mfma_i32_16x16x64_i8 (32768 ops) with NCH chains.
template <int WIDE, int NCH>
__global__ __launch_bounds__(256) void i8_shape_kernel(double* out, int iters) {
  v4i_t ai = {(int)threadIdx.x, 3, 5, 7}, bi = {11, (int)blockIdx.x, 13, 17};
  v16i_t cw[WIDE ? NCH : 1];
  v4i_t cn[WIDE ? 1 : NCH];
#pragma unroll
  for (int j = 0; j < (WIDE ? NCH : 1); ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) cw[j][r] = 0;
#pragma unroll
  for (int j = 0; j < (WIDE ? 1 : NCH); ++j) cn[j] = v4i_t{0, 0, 0, 0};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
      if (WIDE) cw[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai, bi, cw[j], 0, 0, 0);
      else cn[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai, bi, cn[j], 0, 0, 0);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < (WIDE ? NCH : 1); ++j) s += (double)cw[j][0];
#pragma unroll
  for (int j = 0; j < (WIDE ? 1 : NCH); ++j) s += (double)cn[j][0];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

double debug_mfma_peak(int chains, int wps, int iters) {
  if (chains < 1 || (chains > 8 && (chains < 10 || chains > 17)) || chains == 7 || wps < 1 || wps > 8 || iters < 1)
    throw status_error(GWHIP_ERR_ARG, "bad mfma probe args");
  GWHIP_CHECK(hipSetDevice(0));
  hipDeviceProp_t prop;
  GWHIP_CHECK(hipGetDeviceProperties(&prop, 0));
  const int nblocks = prop.multiProcessorCount * wps;  // 4 waves per block -> one per SIMD
  double* out = nullptr;
  GWHIP_CHECK(hipMalloc(&out, (size_t)nblocks * 256 * sizeof(double)));
  auto launch = [&] {
    switch (chains) {
      case 10: hipLaunchKernelGGL(lowp_peak_kernel<0>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 11: hipLaunchKernelGGL(lowp_peak_kernel<1>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 12: hipLaunchKernelGGL(valu_cgemm_probe_kernel, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 13: hipLaunchKernelGGL((i8_shape_kernel<1, 4>), dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 14: hipLaunchKernelGGL((i8_shape_kernel<1, 8>), dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 15: hipLaunchKernelGGL((i8_shape_kernel<0, 16>), dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 16: hipLaunchKernelGGL(mfma4_peak_kernel<4>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 17: hipLaunchKernelGGL(mfma4_peak_kernel<8>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 5: hipLaunchKernelGGL(valu_peak_kernel<1>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 6: hipLaunchKernelGGL(valu_peak_kernel<2>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 1: hipLaunchKernelGGL(mfma_peak_kernel<1>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 2: hipLaunchKernelGGL(mfma_peak_kernel<2>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      case 4: hipLaunchKernelGGL(mfma_peak_kernel<4>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
      default: hipLaunchKernelGGL(mfma_peak_kernel<8>, dim3(nblocks), dim3(256), 0, nullptr, out, iters); break;
    }
  };
  launch();
  GWHIP_CHECK(hipDeviceSynchronize());
  Event e0, e1;
  GWHIP_CHECK(hipEventRecord(e0, nullptr));
  launch();
  GWHIP_CHECK(hipEventRecord(e1, nullptr));
  GWHIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  GWHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipFree(out);
  double per_wave_iter;
  if (chains == 10) per_wave_iter = 8 * 16.0 * 16 * 64 * 2;
  else if (chains == 11) per_wave_iter = 8 * 16.0 * 16 * 32 * 2;
  else if (chains == 12) per_wave_iter = 32 * 64 * 2.0 * 64;  // 32 k x 64 FMA per lane
  else if (chains == 13) per_wave_iter = 4 * 32.0 * 32 * 32 * 2;
  else if (chains == 14) per_wave_iter = 8 * 32.0 * 32 * 32 * 2;
  else if (chains == 15) per_wave_iter = 16 * 16.0 * 16 * 64 * 2;
  else if (chains == 16) per_wave_iter = 4 * 4 * 4.0 * 4 * 4 * 2;
  else if (chains == 17) per_wave_iter = 8 * 4 * 4.0 * 4 * 4 * 2;
  else if (chains == 5) per_wave_iter = 32.0 * 64 * 2;                   // 32 FMA per lane
  else if (chains == 6) per_wave_iter = 32.0 * 64 * 2 + 4 * 2048.0;  // + 4 MFMA
  else per_wave_iter = (chains == 3 ? 4 : (chains > 4 ? 8 : chains)) * 2048.0;
  const double flops = (double)nblocks * 4 * iters * per_wave_iter;
  return flops / (ms * 1e-3) / 1e12;
}

double debug_gemm_bench(int which, int NQ, int nk, int nao, int batch, int shape, int reps) {
  if (NQ <= 0 || nk <= 0 || nao <= 0 || batch <= 0 || reps <= 0 || shape < 0 || shape > 2)
    throw status_error(GWHIP_ERR_ARG, "bad debug_gemm_bench arguments");
  GWHIP_CHECK(hipSetDevice(0));
  const int64_t nao2 = (int64_t)nao * nao, NQ2 = (int64_t)NQ * NQ, vsz = (int64_t)nk * NQ * nao2;
  ZGemmArgs a{};
  int64_t na, nb, nc;
  bool akc = true, bkc = false;
  if (which == 0) {  // Sigma: R[tau] = P[tau] conj(V)
    na = batch * NQ2; nb = vsz; nc = batch * vsz;
    a.M = NQ; a.N = (int)(nk * nao2); a.K = NQ; a.batch = batch; a.kin = NQ; a.nin = (int)nao2;
    a.sAb = NQ2; a.sAm = NQ; a.sAko = 0; a.sAki = 1;
    a.sBb = 0; a.sBko = 0; a.sBki = nao2; a.sBno = NQ * nao2; a.sBni = 1;
    a.sCb = vsz; a.sCm = nao2; a.sCno = NQ * nao2; a.sCni = 1;
  } else {           // P0: P[tau] = Z[tau] conj(V)^T
    na = batch * vsz; nb = vsz; nc = batch * NQ2;
    a.M = NQ; a.N = NQ; a.K = (int)(nk * nao2); a.batch = batch; a.kin = (int)nao2; a.nin = NQ;
    a.sAb = vsz; a.sAm = nao2; a.sAko = NQ * nao2; a.sAki = 1;
    a.sBb = 0; a.sBko = NQ * nao2; a.sBki = 1; a.sBno = 0; a.sBni = nao2;
    a.sCb = NQ2; a.sCm = NQ; a.sCno = 0; a.sCni = 1;
    bkc = true;
  }
  double2 *dA = nullptr, *dB = nullptr, *dC = nullptr;
  GWHIP_CHECK(hipMalloc(&dA, na * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&dB, nb * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&dC, nc * sizeof(double2)));
  hipLaunchKernelGGL(fill_hash_kernel, dim3((unsigned)ceil_div(na, 256)), dim3(256), 0, nullptr, dA, na, 11ull);
  hipLaunchKernelGGL(fill_hash_kernel, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, nullptr, dB, nb, 12ull);
  GWHIP_CHECK(hipGetLastError());
  a.A = dA; a.B = dB; a.C = dC;
  a.alpha = make_double2(1.0, 0.0);
  a.conjB = 1;
  zgemm_f64(a, akc, bkc, nullptr, shape);
  GWHIP_CHECK(hipDeviceSynchronize());
  Event e0, e1;
  GWHIP_CHECK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; ++r) zgemm_f64(a, akc, bkc, nullptr, shape);
  GWHIP_CHECK(hipEventRecord(e1, nullptr));
  GWHIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  GWHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return ms / reps;
}

}  // namespace gwhip

namespace {
template <typename F>
int guarded(F&& f) {
  try {
    f();
    return GWHIP_OK;
  } catch (const gwhip::status_error& e) {
    return e.status;
  } catch (const gwhip::hip_error& e) {
    return e.code == hipErrorOutOfMemory ? GWHIP_ERR_OOM : GWHIP_ERR_DEVICE;
  } catch (const std::exception&) {
    return GWHIP_ERR_DEVICE;
  }
}
}  // namespace

extern "C" {

int gwhip_debug_gemm_bench(int which, int NQ, int nk, int nao, int batch, int shape, int reps, double* ms) {
  if (!ms) return GWHIP_ERR_ARG;
  return guarded([&] { *ms = gwhip::debug_gemm_bench(which, NQ, nk, nao, batch, shape, reps); });
}

int gwhip_debug_mfma_peak(int chains, int waves_per_simd, int iters, double* tflops) {
  if (!tflops) return GWHIP_ERR_ARG;
  return guarded([&] { *tflops = gwhip::debug_mfma_peak(chains, waves_per_simd, iters); });
}

int gwhip_debug_mfma4_layout(double* out) {
  if (!out) return GWHIP_ERR_ARG;
  return guarded([&] {
    GWHIP_CHECK(hipSetDevice(0));
    double* d = nullptr;
    GWHIP_CHECK(hipMalloc(&d, 3 * 64 * sizeof(double)));
    hipLaunchKernelGGL(gwhip::mfma4_layout_kernel, dim3(1), dim3(64), 0, nullptr, d);
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) GWHIP_CHECK(hipMemcpy(out, d, 3 * 64 * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(d);
    GWHIP_CHECK(e);
  });
}

}  // extern "C"
