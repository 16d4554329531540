// Gibbs kernel K = exp(-C/reg) of one cost matrix, stored in the layouts the
// MFMA GEMMs read (every operand k-contiguous): BF16 and FP64, in both
// orientations. Shared by the batched MFMA solver and the serving engine.
#pragma once

#include "common.hpp"

namespace rwig {

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct DenseKernel {
  int M = 0, Nc = 0;
  // padded M, padded Nc (multiples of 64, zero past M and Nc): a BF16 product
  // may run its contraction over the padded length in whole 64-element
  // k-steps (direct-to-LDS tiles) with the same result as over M or Nc
  long ldm = 0, ldn = 0;
  DBuf<uint16_t> Kb, KbT;  // Kb: Kb[i + j*ldm] = K(i,j); KbT: KbT[j + i*ldn] = K(i,j)
  DBuf<double> Kd, KdT;
  // KCT[j + i*ldn] = K(i,j) C(i,j): the transport cost of the batched solver as
  // one FP64 product (built by dense_kernel_cost_build; empty in the serving
  // engine, which needs only the regularised loss)
  DBuf<double> KCT;
};

inline long pad8(long n) { return (n + 7) / 8 * 8; }
inline long pad64(long n) { return (n + 63) / 64 * 64; }

// Device bytes of a DenseKernel of an M x Nc cost.
inline size_t dense_kernel_bytes(int M, int Nc) {
  return (size_t)pad64(M) * Nc * (2 + 8) + (size_t)pad64(Nc) * M * (2 + 8);
}

// Fills dk from the device cost C (M x Nc, leading dimension ldc).
void dense_kernel_build(hipStream_t st, const double *C, long ldc, int M, int Nc, double reg,
                        DenseKernel &dk);

// Fills dk.KCT from the cost dk was built from.
void dense_kernel_cost_build(hipStream_t st, const double *C, long ldc, double reg,
                             DenseKernel &dk);

}  // namespace rwig
