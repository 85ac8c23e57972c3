/*
 * gwhip_debug -- throughput probes for one MI355X (gfx950), in libgwhip_debug.so.
 *
 * Kept out of libgwhip.so: nothing in the GW path calls these.  They measure ceilings (register-only
 * matrix-core and vector-FMA rates) and time the library's FP64 GEMM on operands laid out like one q of
 * the GW contractions.  No reference counterpart.  Status codes are those of gwhip.h.
 */
#ifndef GWHIP_DEBUG_H
#define GWHIP_DEBUG_H

#ifdef __cplusplus
extern "C" {
#endif

/* Time the FP64 GEMM on device 0 with device-resident operands laid out as one q of the
 * Sigma contraction (which = 0: M = NQ, N = nk*nao^2, K = NQ, `batch` tau) or of the P0
 * contraction (which = 1: M = N = NQ, K = nk*nao^2, `batch` tau); `shape` selects the tile
 * (0: 64x64, 1: 128x64, 2: 64x128).  Writes the mean milliseconds per launch over `reps`. */
int gwhip_debug_gemm_bench(int which, int NQ, int nk, int nao, int batch, int shape, int reps, double* ms);

/* Register-only throughput probe on device 0: `chains` independent FP64 MFMA accumulators per wave
 * (1, 2, 4 or 8; 5 = FP64 vector FMA only, 6 = four MFMA chains plus vector FMAs in the same wave,
 * 10 = int8 16x16x64, 11 = bf16 16x16x32, 12 = FP64 vector-FMA complex 4 x 4 outer products with
 * operands from LDS, 13 / 14 = int8 32x32x32 with 4 / 8 chains, 15 = int8 16x16x64 with 16 chains,
 * 16 / 17 = FP64 4x4x4 (four blocks) with 4 / 8 chains),
 * `waves_per_simd` waves on every SIMD, `iters` rounds.  Writes the achieved
 * TFLOP/s (2048 flop per v_mfma_f64_16x16x4f64, 2 per FMA lane). */
int gwhip_debug_mfma_peak(int chains, int waves_per_simd, int iters, double* tflops);

/* Lane layout of v_mfma_f64_4x4x4_4b on device 0: out[3][64] holds D of each lane for three inputs
 * (A = 2^(l%16), B = 2^(16(l%16)); A = l/16 + 1, B = 1; A = 1, B = l/16 + 1). */
int gwhip_debug_mfma4_layout(double* out);

#ifdef __cplusplus
}
#endif
#endif /* GWHIP_DEBUG_H */
