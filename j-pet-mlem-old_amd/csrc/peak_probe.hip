// Measures the gfx950 rates that bound the strip MLEM kernels:
//   - FP32 VALU FMA issue rate (non-packed: packed FP32 is disabled)
//   - v_exp_f32 (hardware exp2) issue rate
//   - global float atomic add rate (L2 atomics), scattered and coalesced
//   - LDS float atomic add rate
// Every loop has a fixed trip count and every index stays inside its buffer.
// Usage: strip_peak_probe           -> JSON objects on stdout
//        strip_peak_probe --copy M  -> copies M MiB once with a plain kernel
//                                     (calibrates HBM byte counters)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#define CHECK(x)                                                        \
  do {                                                                  \
    hipError_t e_ = (x);                                                \
    if (e_ != hipSuccess) {                                             \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));     \
      std::exit(1);                                                     \
    }                                                                   \
  } while (0)

constexpr int kIters = 4096;

__global__ __launch_bounds__(256) void fma_probe(float* out, float s) {
  float a0 = threadIdx.x * 1e-3f, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3;
  float a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
  for (int i = 0; i < kIters; ++i) {
    a0 = __builtin_fmaf(a0, s, 0.5f);
    a1 = __builtin_fmaf(a1, s, 0.5f);
    a2 = __builtin_fmaf(a2, s, 0.5f);
    a3 = __builtin_fmaf(a3, s, 0.5f);
    a4 = __builtin_fmaf(a4, s, 0.5f);
    a5 = __builtin_fmaf(a5, s, 0.5f);
    a6 = __builtin_fmaf(a6, s, 0.5f);
    a7 = __builtin_fmaf(a7, s, 0.5f);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] =
      a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

__global__ __launch_bounds__(256) void exp_probe(float* out, float s) {
  float a0 = -threadIdx.x * 1e-4f, a1 = a0 - 0.1f, a2 = a0 - 0.2f,
        a3 = a0 - 0.3f;
  float a4 = a0 - 0.4f, a5 = a0 - 0.5f, a6 = a0 - 0.6f, a7 = a0 - 0.7f;
  for (int i = 0; i < kIters; ++i) {
    a0 = __builtin_amdgcn_exp2f(a0) * s;
    a1 = __builtin_amdgcn_exp2f(a1) * s;
    a2 = __builtin_amdgcn_exp2f(a2) * s;
    a3 = __builtin_amdgcn_exp2f(a3) * s;
    a4 = __builtin_amdgcn_exp2f(a4) * s;
    a5 = __builtin_amdgcn_exp2f(a5) * s;
    a6 = __builtin_amdgcn_exp2f(a6) * s;
    a7 = __builtin_amdgcn_exp2f(a7) * s;
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] =
      a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

// n_pixels must be a power of two; each lane adds kAtomIters values.
constexpr int kAtomIters = 64;
__global__ __launch_bounds__(256) void atomic_scatter_probe(float* img,
                                                            unsigned mask) {
  unsigned x = (blockIdx.x * blockDim.x + threadIdx.x) * 2654435761u + 12345u;
  for (int i = 0; i < kAtomIters; ++i) {
    x = x * 1664525u + 1013904223u;
    unsafeAtomicAdd(&img[(x >> 4) & mask], 1.0f);
  }
}

// consecutive lanes -> consecutive addresses (one 256 B run per wave)
__global__ __launch_bounds__(256) void atomic_coalesced_probe(float* img,
                                                              unsigned mask) {
  unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  unsigned lane = threadIdx.x & 63;
  unsigned x = wave * 2654435761u + 777u;
  for (int i = 0; i < kAtomIters; ++i) {
    x = x * 1664525u + 1013904223u;
    unsigned base = ((x >> 4) & mask) & ~63u;
    unsafeAtomicAdd(&img[base + lane], 1.0f);
  }
}

constexpr int kLdsWords = 16384;  // 64 KiB
__global__ __launch_bounds__(256) void lds_atomic_probe(float* out) {
  __shared__ float buf[kLdsWords];
  for (int i = threadIdx.x; i < kLdsWords; i += blockDim.x) buf[i] = 0.f;
  __syncthreads();
  unsigned x = (blockIdx.x * blockDim.x + threadIdx.x) * 2654435761u + 99u;
  for (int i = 0; i < kIters / 4; ++i) {
    x = x * 1664525u + 1013904223u;
    atomicAdd(&buf[(x >> 8) & (kLdsWords - 1)], 1.0f);
  }
  __syncthreads();
  float s = 0.f;
  for (int i = threadIdx.x; i < kLdsWords; i += blockDim.x) s += buf[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// LDS access patterns over a 32 KiB buffer, kLdsIters accesses per lane.
//   mode 0: ds_read_b32, lane i -> word (base + i)        (conflict-free)
//   mode 1: ds_add_f32,  lane i -> word (base + i)        (conflict-free)
//   mode 2: ds_add_f32,  lane i -> word (col*64 + i), col varying per lane
//           (column-major band: one row per lane)
//   mode 3: ds_add_f32,  random word
//   mode 4: plain read-modify-write, lane i -> word (base + i), each wave in
//           its own 4 KiB slice (no two lanes or waves share a word)
constexpr int kLdsIters = 512;
constexpr int kLdsWords2 = 8192;
__global__ __launch_bounds__(256) void lds_pattern_probe(float* out, int mode) {
  __shared__ float buf[kLdsWords2];
  for (int i = threadIdx.x; i < kLdsWords2; i += blockDim.x) buf[i] = 0.f;
  __syncthreads();
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned x = (blockIdx.x * blockDim.x + threadIdx.x) * 2654435761u + 7u;
  float acc = 0.f;
  for (int i = 0; i < kLdsIters; ++i) {
    x = x * 1664525u + 1013904223u;
    unsigned base = (x >> 20) & 63u;  // same for... varies per lane
    if (mode == 0) {
      unsigned b = (unsigned)(i * 64) & (kLdsWords2 - 1);
      acc += buf[b + lane];
    } else if (mode == 1) {
      unsigned b = (unsigned)(i * 64) & (kLdsWords2 - 1);
      atomicAdd(&buf[b + lane], 1.0f);
    } else if (mode == 2) {
      unsigned col = (base + (unsigned)i) & 127u;
      atomicAdd(&buf[col * 64 + lane], 1.0f);
    } else if (mode == 3) {
      atomicAdd(&buf[(x >> 8) & (kLdsWords2 - 1)], 1.0f);
    } else {
      unsigned b = wave * 1024 + (((unsigned)i * 64) & 1023u);
      buf[b + lane] = buf[b + lane] + 1.0f;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kLdsWords2; i += blockDim.x) acc += buf[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void copy_kernel(float4* __restrict__ dst,
                                                   const float4* __restrict__ src,
                                                   size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int copy_mode(size_t mib) {
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 1, bytes));
  CHECK(hipDeviceSynchronize());
  copy_kernel<<<4096, 256>>>(b, a, n);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  std::printf("{\"copy_bytes\": %zu, \"read_kib\": %zu, \"write_kib\": %zu}\n",
              bytes, bytes >> 10, bytes >> 10);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}

template <typename F>
float time_ms(F launch, int reps) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  launch();
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(a));
  for (int r = 0; r < reps; ++r) launch();
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, a, b));
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  return ms / reps;
}

int main(int argc, char** argv) {
  if (argc == 3 && std::string(argv[1]) == "--copy")
    return copy_mode((size_t)std::atol(argv[2]));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const int blocks = cus * 16;
  const int threads = 256;
  float* out = nullptr;
  CHECK(hipMalloc(&out, sizeof(float) * blocks * threads));
  const unsigned n_img = 1u << 20;  // 4 MiB image
  float* img = nullptr;
  CHECK(hipMalloc(&img, sizeof(float) * n_img));
  CHECK(hipMemset(img, 0, sizeof(float) * n_img));

  const double lanes = (double)blocks * threads;
  float t_fma = time_ms([&] { fma_probe<<<blocks, threads>>>(out, 0.999f); }, 5);
  double fma_rate = lanes * kIters * 8 / (t_fma * 1e-3);  // FMA per second
  float t_exp = time_ms([&] { exp_probe<<<blocks, threads>>>(out, 0.999f); }, 5);
  double exp_rate = lanes * kIters * 8 / (t_exp * 1e-3);
  float t_as = time_ms(
      [&] { atomic_scatter_probe<<<blocks, threads>>>(img, n_img - 1); }, 5);
  double as_rate = lanes * kAtomIters / (t_as * 1e-3);
  float t_ac = time_ms(
      [&] { atomic_coalesced_probe<<<blocks, threads>>>(img, n_img - 1); }, 5);
  double ac_rate = lanes * kAtomIters / (t_ac * 1e-3);
  float t_lds = time_ms([&] { lds_atomic_probe<<<blocks, threads>>>(out); }, 5);
  double lds_rate = lanes * (kIters / 4) / (t_lds * 1e-3);
  double lds_pattern[5];
  for (int mode = 0; mode < 5; ++mode) {
    float t = time_ms([&] { lds_pattern_probe<<<blocks, threads>>>(out, mode); }, 5);
    lds_pattern[mode] = lanes * kLdsIters / (t * 1e-3);
  }
  CHECK(hipDeviceSynchronize());
  std::printf("{\"lds_read_conflict_free_per_s\": %.4e, "
              "\"lds_atomic_conflict_free_per_s\": %.4e, "
              "\"lds_atomic_column_major_per_s\": %.4e, "
              "\"lds_atomic_random_per_s\": %.4e, "
              "\"lds_rmw_private_per_s\": %.4e}\n",
              lds_pattern[0], lds_pattern[1], lds_pattern[2], lds_pattern[3],
              lds_pattern[4]);
  std::printf(
      "{\"device\": \"%s\", \"arch\": \"%s\", \"cus\": %d, \"clock_khz\": %d, "
      "\"fp32_fma_per_s\": %.4e, \"fp32_tflops_nonpacked\": %.2f, "
      "\"exp2_per_s\": %.4e, \"global_atomic_scattered_per_s\": %.4e, "
      "\"global_atomic_coalesced_per_s\": %.4e, \"lds_atomic_per_s\": %.4e, "
      "\"lds_bytes_per_block\": %zu}\n",
      prop.name, prop.gcnArchName, cus, prop.clockRate, fma_rate,
      2 * fma_rate * 1e-12, exp_rate, as_rate, ac_rate, lds_rate,
      (size_t)prop.sharedMemPerBlock);
  CHECK(hipFree(out));
  CHECK(hipFree(img));
  return 0;
}
