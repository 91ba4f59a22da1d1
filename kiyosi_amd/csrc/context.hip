// Context lifecycle, the ordered partial merge, flags, timing and the host
// side of the scrambled Sobol setup.
#include <algorithm>
#include <cstring>

#include "host_common.hpp"

extern "C" const uint32_t kg_joe_kuo_words[];
extern "C" const uint32_t kg_joe_kuo_end[];

namespace kg {

// Merges per-block partials [block][slot] for `slots` independent Moments,
// one workgroup per slot (merge_run: the parallel form of the reference's
// merge, monte_carlo.rs:1800-1809 and :2192-2222). The result depends only
// on the partials, not on scheduling, so a seed replays exactly. With
// `flags`, workgroup 0 also copies the call's flag word behind the slots, so
// one copy brings back both.
__global__ void __launch_bounds__(MERGE_THREADS) merge_partials(const Moments* __restrict__ partials,
                                                                int64_t blocks, int slots,
                                                                Moments* __restrict__ out,
                                                                const uint32_t* __restrict__ flags) {
  __shared__ double smem[4 * MERGE_THREADS / WAVE];
  const int s = blockIdx.x;
  if (flags != nullptr && s == 0 && threadIdx.x == 0) *(uint32_t*)(out + slots) = *flags;
  const Moments m = merge_run(partials + s, slots, blocks, smem);
  if (threadIdx.x == 0) out[s] = m;
}

__global__ void clear_word(uint32_t* w) { *w = 0u; }

uint32_t* reset_flags(kg_context* ctx) {
  uint32_t* f = (uint32_t*)ctx->flags.reserve(256);
  // A kernel rather than hipMemsetAsync keeps the sequence capturable.
  hipLaunchKernelGGL(clear_word, dim3(1), dim3(1), 0, ctx->stream, f);
  KG_HIP(hipGetLastError());
  return f;
}

void check_flags(kg_context* ctx) {
  uint32_t h = 0;
  KG_HIP(hipMemcpyAsync(&h, ctx->flags.ptr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP(hipStreamSynchronize(ctx->stream));
  if (h & FLAG_CANCELLED) fail(KG_INTERRUPTED, "operation interrupted");
  if (h & FLAG_NONFINITE) unstable();
}

constexpr size_t PINNED_RESULT_BYTES = 4096;

static void reserve_pinned(kg_context* ctx, size_t bytes) {
  if (bytes <= ctx->pinned_bytes) return;
  KG_HIP(hipStreamSynchronize(ctx->stream));  // no copy may still read the old buffer
  if (ctx->pinned) KG_HIP(hipHostFree(ctx->pinned));
  ctx->pinned = nullptr;
  ctx->pinned_bytes = 0;
  const size_t size = std::max<size_t>(bytes, size_t(1) << 20);
  KG_HIP(hipHostMalloc(&ctx->pinned, size, hipHostMallocDefault));
  ctx->pinned_bytes = size;
}

void merge_and_fetch(kg_context* ctx, const Moments* partials, int64_t blocks, int slots,
                     Moments* host_out) {
  const size_t bytes = sizeof(Moments) * slots;
  if (bytes <= PINNED_RESULT_BYTES) {
    // Merged straight into the pinned host buffer, as finish_sampling does.
    reserve_pinned(ctx, PINNED_RESULT_BYTES);
    hipLaunchKernelGGL(merge_partials, dim3(slots), dim3(MERGE_THREADS), 0, ctx->stream, partials, blocks, slots,
                       (Moments*)ctx->pinned, (const uint32_t*)nullptr);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(host_out, ctx->pinned, bytes);
    return;
  }
  Moments* d_out = (Moments*)ctx->results.reserve(bytes);
  hipLaunchKernelGGL(merge_partials, dim3(slots), dim3(MERGE_THREADS), 0, ctx->stream, partials, blocks, slots,
                     d_out, (const uint32_t*)nullptr);
  KG_HIP(hipGetLastError());
  KG_HIP(hipMemcpyAsync(host_out, d_out, bytes, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP(hipStreamSynchronize(ctx->stream));
}

void upload_async(kg_context* ctx, void* dst, const void* src, size_t n) {
  if (n == 0) return;
  const size_t at = PINNED_RESULT_BYTES + ((ctx->staging_used + 255) & ~size_t(255));
  if (at + n > ctx->pinned_bytes) {
    // The staged bytes restart at zero, so every copy still reading them
    // must finish first; the buffer grows if one upload does not fit.
    KG_HIP(hipStreamSynchronize(ctx->stream));
    reserve_pinned(ctx, 2 * (PINNED_RESULT_BYTES + n));
    ctx->staging_used = 0;
    upload_async(ctx, dst, src, n);
    return;
  }
  unsigned char* stage = (unsigned char*)ctx->pinned + at;
  std::memcpy(stage, src, n);
  KG_HIP(hipMemcpyAsync(dst, stage, n, hipMemcpyHostToDevice, ctx->stream));
  ctx->staging_used = at + n - PINNED_RESULT_BYTES;
}

void finish_sampling(kg_context* ctx, const Moments* partials, int64_t blocks, int slots, Moments* host_out,
                     double* device_ms) {
  KG_HIP(hipEventRecord(ctx->stop, ctx->stream));
  reserve_pinned(ctx, PINNED_RESULT_BYTES);
  const size_t result_bytes = sizeof(Moments) * slots;
  if (result_bytes + sizeof(uint32_t) > PINNED_RESULT_BYTES) fail(KG_INVALID_INPUT, "too many merged outputs");
  // merge_partials writes the merged slots, then the flag word, straight
  // into the pinned host buffer (device-visible), so no copy follows it.
  unsigned char* pin = (unsigned char*)ctx->pinned;
  hipLaunchKernelGGL(merge_partials, dim3(slots), dim3(MERGE_THREADS), 0, ctx->stream, partials, blocks, slots,
                     (Moments*)pin, (const uint32_t*)ctx->flags.ptr);
  KG_HIP(hipGetLastError());
  KG_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t h;
  std::memcpy(&h, pin + result_bytes, sizeof(h));
  if (h & FLAG_CANCELLED) fail(KG_INTERRUPTED, "operation interrupted");
  if (h & FLAG_NONFINITE) unstable();
  std::memcpy(host_out, pin, result_bytes);
  float ms = 0.f;
  KG_HIP(hipEventElapsedTime(&ms, ctx->start, ctx->stop));
  *device_ms = ms;
}

void finish_sampling_into(kg_context* ctx, const Moments* partials, int64_t blocks, int slots, Moments* d_out,
                          const uint32_t* d_flags, Moments* host_out, double* device_ms) {
  KG_HIP(hipEventRecord(ctx->stop, ctx->stream));
  hipLaunchKernelGGL(merge_partials, dim3(slots), dim3(MERGE_THREADS), 0, ctx->stream, partials, blocks, slots,
                     d_out, d_flags);
  KG_HIP(hipGetLastError());
  // Merged slots and flag word into a fixed host buffer: pageable memory,
  // copied before the synchronisation returns.
  constexpr int MAX_SLOTS = 4;
  if (slots > MAX_SLOTS) fail(KG_INVALID_INPUT, "too many merged outputs");
  unsigned char host[MAX_SLOTS * sizeof(Moments) + sizeof(uint32_t)];
  const size_t result_bytes = sizeof(Moments) * slots;
  KG_HIP(hipMemcpyAsync(host, d_out, result_bytes + sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t h;
  std::memcpy(&h, host + result_bytes, sizeof(h));
  if (h & FLAG_CANCELLED) fail(KG_INTERRUPTED, "operation interrupted");
  if (h & FLAG_NONFINITE) unstable();
  std::memcpy(host_out, host, result_bytes);
  float ms = 0.f;
  KG_HIP(hipEventElapsedTime(&ms, ctx->start, ctx->stop));
  *device_ms = ms;
}

void clear_flag_word(kg_context* ctx, uint32_t* word) {
  hipLaunchKernelGGL(clear_word, dim3(1), dim3(1), 0, ctx->stream, word);
  KG_HIP(hipGetLastError());
}

void timing_begin(kg_context* ctx) { KG_HIP(hipEventRecord(ctx->start, ctx->stream)); }
double timing_end(kg_context* ctx) {
  KG_HIP(hipEventRecord(ctx->stop, ctx->stream));
  KG_HIP(hipEventSynchronize(ctx->stop));
  float ms = 0.f;
  KG_HIP(hipEventElapsedTime(&ms, ctx->start, ctx->stop));
  return ms;
}

// Linear matrix scrambling and the digital shift of one dimension
// (sobol.rs:100-133): Pcg32(seed, dimension) draws the shift, then one
// lower-triangular row per bit with a unit diagonal; column c of the result
// goes to out[c * stride].
static void scramble_dimension(const uint64_t (&raw)[52], uint64_t seed, uint32_t d, uint64_t* out,
                               size_t stride, uint64_t* shift) {
  constexpr int B = 52;
  constexpr uint64_t MASK = (1ULL << B) - 1;
  Pcg32 g;
  g.init(seed, d);
  auto word = [&g]() {
    uint64_t hi = g.next();
    return (hi << 32) | (uint64_t)g.next();
  };
  *shift = word() & MASK;
  uint64_t rows[B];
  for (int row = 0; row < B; ++row) {
    const int bit = B - 1 - row;
    rows[row] = (word() & (MASK << (bit + 1))) | (1ULL << bit);
  }
  for (int c = 0; c < B; ++c) {
    uint64_t v = 0;
    for (int row = 0; row < B; ++row)
      v |= (uint64_t)(__builtin_popcountll(raw[c] & rows[row]) & 1) << (B - 1 - row);
    out[(size_t)c * stride] = v;
  }
}

void sobol_dimension_zero(uint64_t seed, uint64_t (&dirs)[52], uint64_t* shift) {
  uint64_t raw[52];
  for (int c = 0; c < 52; ++c) raw[c] = 1ULL << (51 - c);
  scramble_dimension(raw, seed, 0, dirs, 1, shift);
}

// The unscrambled 52-bit direction integers of dimensions [first, dims) from
// the Joe-Kuo records (sobol.rs:85-99): dimension 0 is the identity,
// others follow the primitive-polynomial recurrence from their initial
// values. raw[(d - first) * 52 + c]; `pos` is the record offset of `first`.
static void raw_directions(const std::vector<uint32_t>& table, uint32_t first, uint32_t dims, size_t* pos,
                           std::vector<uint64_t>& raw) {
  constexpr int B = 52;
  raw.assign((size_t)(dims - first) * B, 0);
  for (uint32_t d = first; d < dims; ++d) {
    uint64_t* r = &raw[(size_t)(d - first) * B];
    if (d == 0) {
      for (int c = 0; c < B; ++c) r[c] = 1ULL << (B - 1 - c);
      continue;
    }
    if (*pos + 2 > table.size()) fail(KG_INVALID_INPUT, "Sobol direction table is truncated");
    const int degree = (int)table[*pos];
    const uint32_t poly = table[*pos + 1];
    for (int c = 0; c < degree && c < B; ++c) r[c] = (uint64_t)table[*pos + 2 + c] << (B - 1 - c);
    *pos += (size_t)degree + 2;
    for (int c = degree; c < B; ++c) {
      uint64_t v = r[c - degree] ^ (r[c - degree] >> degree);
      for (int k = 1; k < degree; ++k)
        if ((poly >> (degree - 1 - k)) & 1u) v ^= r[c - k];
      r[c] = v;
    }
  }
}

// SobolSequence::new (engines/monte_carlo/sobol.rs:85-134) on the host:
// raw directions, then linear matrix scrambling and a digital shift drawn
// from Pcg32(seed, dimension).
static void build_sobol(const std::vector<uint32_t>& table, uint32_t dims, uint64_t seed,
                        std::vector<uint64_t>& dirs, std::vector<uint64_t>& shifts) {
  constexpr int B = 52;
  dirs.assign((size_t)dims * B, 0);
  shifts.assign(dims, 0);
  std::vector<uint64_t> raw;
  size_t pos = 0;
  raw_directions(table, 0, dims, &pos, raw);
  for (uint32_t d = 0; d < dims; ++d) {
    uint64_t r[B];
    for (int c = 0; c < B; ++c) r[c] = raw[(size_t)d * B + c];
    scramble_dimension(r, seed, d, &dirs[d], dims, &shifts[d]);
  }
}

// The same scrambling on the device, one wave per dimension: lane r draws
// the scramble row of bit 51 - r (PCG words 1 + r of Pcg32(seed, d), i.e.
// outputs 2 + 2r and 3 + 2r, reached by skip-ahead), lane 52 the shift (word
// 0); then lane c forms scrambled column c as the parities of raw column c
// against the 52 rows.
__global__ void __launch_bounds__(256) sobol_scramble_kernel(const uint64_t* __restrict__ raw, uint32_t dims,
                                                             uint64_t seed, uint64_t* __restrict__ dirs,
                                                             uint64_t* __restrict__ shifts) {
  constexpr int B = 52;
  constexpr uint64_t MASK = (1ULL << B) - 1;
  __shared__ uint64_t rows[4][B];
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const uint32_t d = blockIdx.x * 4 + wave;
  const bool live = d < dims;
  if (live && lane <= B) {
    Pcg32 g;
    g.init(seed, d);
    const int word = lane == B ? 0 : 1 + lane;
    g.advance(2 * (uint64_t)word);
    const uint64_t hi = g.next();
    const uint64_t w = (hi << 32) | (uint64_t)g.next();
    if (lane == B) {
      shifts[d] = w & MASK;
    } else {
      const int bit = B - 1 - lane;
      rows[wave][lane] = (w & (MASK << (bit + 1))) | (1ULL << bit);
    }
  }
  __syncthreads();
  if (live && lane < B) {
    const uint64_t col = raw[(size_t)d * B + lane];
    uint64_t v = 0;
    for (int row = 0; row < B; ++row)
      v |= (uint64_t)(__builtin_popcountll(col & rows[wave][row]) & 1) << (B - 1 - row);
    dirs[(size_t)lane * dims + d] = v;
  }
}

void sobol_tables(kg_context* ctx, uint32_t dims, uint64_t seed, std::vector<uint64_t>& dirs,
                  std::vector<uint64_t>& shifts) {
  if (ctx->joe_kuo.empty()) ctx->joe_kuo.assign(kg_joe_kuo_words, kg_joe_kuo_end);
  build_sobol(ctx->joe_kuo, dims, seed, dirs, shifts);
}

// The raw directions of the first `dims` dimensions on the device, kept
// across calls and seeds (they do not depend on the seed); grown on demand.
static const uint64_t* raw_on_device(kg_context* ctx, uint32_t dims) {
  if (ctx->sobol_raw_dims >= dims) return (const uint64_t*)ctx->sobol_raw.ptr;
  if (ctx->joe_kuo.empty()) ctx->joe_kuo.assign(kg_joe_kuo_words, kg_joe_kuo_end);
  std::vector<uint64_t> raw;
  size_t pos = 0;
  raw_directions(ctx->joe_kuo, 0, dims, &pos, raw);
  KG_HIP(hipStreamSynchronize(ctx->stream));  // no kernel may still read the old buffer
  void* d = ctx->sobol_raw.reserve(raw.size() * sizeof(uint64_t));
  KG_HIP(hipMemcpy(d, raw.data(), raw.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  ctx->sobol_raw_dims = dims;
  return (const uint64_t*)d;
}

const SobolSetup& sobol_setup(kg_context* ctx, uint32_t dims, uint64_t seed) {
  SobolSetup& s = ctx->sobol;
  if (s.valid && s.dims == dims && s.seed == seed) return s;
  s.valid = false;
  const uint64_t* raw = raw_on_device(ctx, dims);
  uint64_t* dd = (uint64_t*)s.d_directions.reserve((size_t)dims * 52 * sizeof(uint64_t));
  uint64_t* ds = (uint64_t*)s.d_shifts.reserve((size_t)dims * sizeof(uint64_t));
  hipLaunchKernelGGL(sobol_scramble_kernel, dim3((dims + 3) / 4), dim3(256), 0, ctx->stream, raw, dims, seed, dd,
                     ds);
  KG_HIP(hipGetLastError());
  s.dims = dims;
  s.seed = seed;
  s.valid = true;
  return s;
}

}  // namespace kg

extern "C" {

const char* kg_version(void) { return "kiyosi_amd 0.1.0 (gfx950)"; }

static std::mutex g_setup_mu;  // per-device one-time set-up across contexts

kg_status kg_context_create(int32_t device, kg_context** out) {
  if (!out) return KG_INVALID_INPUT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return KG_DEVICE_ERROR;
  if (device < 0 || device >= count) return KG_INVALID_INPUT;
  kg_context* ctx = new (std::nothrow) kg_context();
  if (!ctx) return KG_COMPUTATION_FAILURE;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipEventCreate(&ctx->start) != hipSuccess || hipEventCreate(&ctx->stop) != hipSuccess) {
    kg_context_destroy(ctx);
    return KG_DEVICE_ERROR;
  }
  try {
    std::lock_guard<std::mutex> lock(g_setup_mu);
    kg::noalloc_setup(device);
  } catch (const kg::Error&) {
    kg_context_destroy(ctx);
    return KG_DEVICE_ERROR;
  }
  *out = ctx;
  return KG_OK;
}

void kg_context_destroy(kg_context* ctx) {
  if (!ctx) return;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->partials.release();
    ctx->scratch.release();
    ctx->flags.release();
    ctx->results.release();
    ctx->sobol.d_directions.release();
    ctx->sobol.d_shifts.release();
    ctx->sobol_raw.release();
    for (kg::DeviceBuffer* b : {&ctx->book_blob, &ctx->book_items, &ctx->book_tiles, &ctx->book_parts,
                                &ctx->book_out, &ctx->book_counter, &ctx->policy, &ctx->lsmc_tables,
                                &ctx->mlp_state})
      b->release();
    if (ctx->book_graph) (void)hipGraphExecDestroy(ctx->book_graph);
    ctx->book_graph = nullptr;
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    if (ctx->start) (void)hipEventDestroy(ctx->start);
    if (ctx->stop) (void)hipEventDestroy(ctx->stop);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

const char* kg_last_error(const kg_context* ctx) { return ctx ? ctx->error.c_str() : ""; }

kg_status kg_cancel_token_create(kg_cancel_token** out) {
  if (!out) return KG_INVALID_INPUT;
  *out = nullptr;
  kg_cancel_token* t = new (std::nothrow) kg_cancel_token();
  if (!t) return KG_COMPUTATION_FAILURE;
  void* host = nullptr;
  // Coherent, so a kernel reading the word sees a raise made while it runs.
  if (hipHostMalloc(&host, 256, hipHostMallocMapped | hipHostMallocPortable | hipHostMallocCoherent) !=
      hipSuccess) {
    delete t;
    return KG_DEVICE_ERROR;
  }
  void* dev = nullptr;
  if (hipHostGetDevicePointer(&dev, host, 0) != hipSuccess) {
    (void)hipHostFree(host);
    delete t;
    return KG_DEVICE_ERROR;
  }
  t->host = (volatile int32_t*)host;
  *t->host = 0;
  t->device = (int32_t*)dev;
  *out = t;
  return KG_OK;
}

void kg_cancel_token_destroy(kg_cancel_token* token) {
  if (!token) return;
  if (token->host) (void)hipHostFree((void*)token->host);
  delete token;
}

void kg_cancel(kg_cancel_token* token) {
  if (token) __atomic_store_n(token->host, 1, __ATOMIC_RELEASE);
}

void kg_cancel_reset(kg_cancel_token* token) {
  if (token) __atomic_store_n(token->host, 0, __ATOMIC_RELEASE);
}

int32_t kg_cancel_requested(const kg_cancel_token* token) {
  return token ? __atomic_load_n(token->host, __ATOMIC_ACQUIRE) : 0;
}

kg_status kg_context_set_cancel_token(kg_context* ctx, kg_cancel_token* token) {
  if (!ctx) return KG_INVALID_INPUT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->cancel = token;
  return KG_OK;
}
}
