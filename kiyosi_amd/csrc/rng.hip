// Generator diagnostics: PCG32 outputs, block-stream normals and scrambled
// Sobol points computed on the GPU with per-thread skip-ahead, for bit-level
// comparison with the reference streams (monte_carlo.rs:2224-2277,
// 1678-1703; engines/monte_carlo/sobol.rs:136-243), and the single-precision
// inverse normal of the FP32-step kernels on given 32-bit inputs.
#include "fp32_math.hpp"
#include "host_common.hpp"

namespace kg {

__global__ void pcg32_outputs_kernel(uint64_t seed, uint64_t stream, uint64_t first,
                                     uint64_t count, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  Pcg32 g;
  g.init(seed, stream);
  g.advance(first + i);
  out[i] = g.next();
}

// Thread i produces normal first+i of the stream, from Box-Muller pair
// (first+i)/2 located by skip-ahead.
__global__ void pcg_normals_kernel(uint64_t seed, uint64_t stream, uint64_t first, uint64_t count,
                                   double* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  NormalStream s;
  s.start(seed, stream, first + i);
  out[i] = s.next();
}

// The FP32-step kernels' inverse normal of (bits + 1/2) / 2^32.
__global__ void fp32_normals_kernel(const uint32_t* __restrict__ bits, uint64_t count, float* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = f32::normal_from_bits(bits[i]);
}

__global__ void sobol_points_kernel(const uint64_t* __restrict__ dirs,
                                    const uint64_t* __restrict__ shifts, uint32_t dims,
                                    uint64_t first, uint64_t count, uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count * dims) return;
  uint64_t p = first + i / dims;
  uint32_t d = (uint32_t)(i % dims);
  uint64_t gray = p ^ (p >> 1);
  uint64_t v = shifts[d];
  while (gray) {
    int c = __builtin_ctzll(gray);
    v ^= dirs[(uint64_t)c * dims + d];
    gray &= gray - 1;
  }
  out[i] = v;
}

static void check_count(uint64_t count) {
  if (count > (uint64_t(1) << 32)) fail(KG_INVALID_INPUT, "count is limited to 2^32");
}

}  // namespace kg

using namespace kg;

extern "C" {

kg_status kg_pcg32_outputs(kg_context* ctx, uint64_t seed, uint64_t stream, uint64_t first,
                           uint64_t count, uint32_t* out) {
  return guarded(ctx, [&] {
    check_count(count);
    if (count == 0) return;
    if (!out) fail(KG_INVALID_INPUT, "out must not be null");
    uint32_t* d = (uint32_t*)ctx->scratch.reserve(count * sizeof(uint32_t));
    hipLaunchKernelGGL(pcg32_outputs_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream,
                       seed, stream, first, count, d);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemcpyAsync(out, d, count * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kg_status kg_fp32_normals_from_bits(kg_context* ctx, const uint32_t* bits, uint64_t count, float* out) {
  return guarded(ctx, [&] {
    check_count(count);
    if (count == 0) return;
    if (!bits || !out) fail(KG_INVALID_INPUT, "bits and out must not be null");
    uint32_t* d_in = (uint32_t*)ctx->scratch.reserve(count * (sizeof(uint32_t) + sizeof(float)));
    float* d_out = (float*)(d_in + count);
    KG_HIP(hipMemcpyAsync(d_in, bits, count * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(fp32_normals_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, d_in, count,
                       d_out);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemcpyAsync(out, d_out, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kg_status kg_pcg_normals(kg_context* ctx, uint64_t seed, uint64_t stream, uint64_t first,
                         uint64_t count, double* out) {
  return guarded(ctx, [&] {
    check_count(count);
    if (count == 0) return;
    if (!out) fail(KG_INVALID_INPUT, "out must not be null");
    double* d = (double*)ctx->scratch.reserve(count * sizeof(double));
    hipLaunchKernelGGL(pcg_normals_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream,
                       seed, stream, first, count, d);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemcpyAsync(out, d, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kg_status kg_sobol_points(kg_context* ctx, uint32_t dims, uint64_t seed, uint64_t first,
                          uint64_t count, uint64_t* out) {
  return guarded(ctx, [&] {
    if (dims < 1 || dims > 21201) fail(KG_INVALID_INPUT, "Sobol dimension must be in 1..=21,201");
    check_count(count * dims);
    if (count == 0) return;
    if (!out) fail(KG_INVALID_INPUT, "out must not be null");
    const SobolSetup& s = sobol_setup(ctx, dims, seed);
    uint64_t n = count * dims;
    uint64_t* d = (uint64_t*)ctx->scratch.reserve(n * sizeof(uint64_t));
    hipLaunchKernelGGL(sobol_points_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       (const uint64_t*)s.d_directions.ptr, (const uint64_t*)s.d_shifts.ptr, dims,
                       first, count, d);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemcpyAsync(out, d, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}
}
