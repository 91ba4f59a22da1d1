// Scrambled Sobol points on the device, bit-identical to SobolCursor
// (engines/monte_carlo/sobol.rs:136-243).
//
// Point n of dimension d is shift_d XOR the directions of the set bits of
// gray(n) = n ^ (n >> 1). For n = 256*b + j (j < 256) the Gray code splits
// as gray(256*b) ^ gray(j), so a workgroup that owns sample block b stages
// two tables in LDS:
//   lo[d][h] = XOR of directions 0..3 selected by h          (block-independent)
//   hi[d][h] = base_d ^ XOR of directions 4..7 selected by h (base_d = point 256*b)
// and a lane reads point j as lo[d][g & 15] ^ hi[d][g >> 4], g = gray(j).
// The FP32 kernels stage the top 32 of the 52 bits only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kg {

constexpr size_t SOBOL_LDS_LIMIT = 64 * 1024;

// Both tables plus one dimension of slack after `hi`, so that a read one
// dimension past the path (a loop's last look-ahead, whose value goes
// unused) stays inside the allocation.
inline size_t sobol_lds_bytes(int dims, bool fp32) {
  return ((size_t)dims * 32 + 16) * (fp32 ? sizeof(uint32_t) : sizeof(uint64_t));
}

// Dimension d of the point at index `first`, a multiple of 256: shift_d
// XOR the directions of the set bits of gray(first).
// The set bits are taken four at a time, so that their loads are in flight
// together rather than one after another.
__device__ inline uint64_t sobol_block_base(const uint64_t* __restrict__ dirs,
                                            const uint64_t* __restrict__ shifts, int dims, uint64_t first,
                                            int d) {
  uint64_t base = shifts[d];
  uint64_t g = first ^ (first >> 1);
  while (g) {
    const uint64_t g1 = g & (g - 1), g2 = g1 & (g1 - 1), g3 = g2 & (g2 - 1);
    const uint64_t a = dirs[(uint64_t)__builtin_ctzll(g) * dims + d];
    const uint64_t b = g1 ? dirs[(uint64_t)__builtin_ctzll(g1) * dims + d] : 0;
    const uint64_t c = g2 ? dirs[(uint64_t)__builtin_ctzll(g2) * dims + d] : 0;
    const uint64_t e = g3 ? dirs[(uint64_t)__builtin_ctzll(g3) * dims + d] : 0;
    base ^= a ^ b ^ c ^ e;
    g = g3 & (g3 - 1);
  }
  return base;
}

// Row d of both tables: lo[h] = XOR of directions 0..3 selected by h,
// hi[h] = base XOR directions 4..7 selected by h; stored >> SHIFT.
template <typename W, int SHIFT>
__device__ inline void sobol_table_rows(const uint64_t* __restrict__ dirs, int dims, int d, uint64_t base,
                                        W* lo, W* hi) {
  uint64_t v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = dirs[(uint64_t)k * dims + d];
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    uint64_t a = 0, b = base;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (h & (1 << k)) {
        a ^= v[k];
        b ^= v[4 + k];
      }
    lo[h] = (W)(a >> SHIFT);
    hi[h] = (W)(b >> SHIFT);
  }
}

template <typename W, int SHIFT>
__device__ inline void sobol_build_lds(const uint64_t* __restrict__ dirs,
                                       const uint64_t* __restrict__ shifts, int dims,
                                       uint64_t first, unsigned char* lds, const void** lo_out,
                                       const void** hi_out) {
  W* lo = (W*)lds;
  W* hi = lo + (size_t)dims * 16;
  for (int d = threadIdx.x; d < dims; d += blockDim.x)
    sobol_table_rows<W, SHIFT>(dirs, dims, d, sobol_block_base(dirs, shifts, dims, first, d), lo + d * 16,
                               hi + d * 16);
  __syncthreads();
  *lo_out = lo;
  *hi_out = hi;
}

__device__ inline void sobol_build_lds64(const uint64_t* dirs, const uint64_t* shifts, int dims,
                                         uint64_t first, unsigned char* lds, const void** lo,
                                         const void** hi) {
  sobol_build_lds<uint64_t, 0>(dirs, shifts, dims, first, lds, lo, hi);
}
__device__ inline void sobol_build_lds32(const uint64_t* dirs, const uint64_t* shifts, int dims,
                                         uint64_t first, unsigned char* lds, const void** lo,
                                         const void** hi) {
  sobol_build_lds<uint32_t, 20>(dirs, shifts, dims, first, lds, lo, hi);
}

__device__ inline uint64_t sobol_lds_bits64(const uint64_t* lo, const uint64_t* hi, int d,
                                            uint32_t glo, uint32_t ghi) {
  return lo[d * 16 + glo] ^ hi[d * 16 + ghi];
}
__device__ inline uint32_t sobol_lds_bits32(const uint32_t* lo, const uint32_t* hi, int d,
                                            uint32_t glo, uint32_t ghi) {
  return lo[d * 16 + glo] ^ hi[d * 16 + ghi];
}

// Fallback for paths too long for the LDS tables: the point from global
// directions, one XOR per set Gray-code bit.
__device__ inline uint64_t sobol_point_direct(const uint64_t* __restrict__ dirs,
                                              const uint64_t* __restrict__ shifts, int dims,
                                              uint64_t point, int d) {
  uint64_t g = point ^ (point >> 1);
  uint64_t v = shifts[d];
  while (g) {
    v ^= dirs[(uint64_t)__builtin_ctzll(g) * dims + d];
    g &= g - 1;
  }
  return v;
}

}  // namespace kg
