// ctas-peak-probe: measures the peaks the backprojection roofline is priced
// against on the device it runs on: the shader clock, LDS read bandwidth
// (ds_read_b32 / b64 / b128 with conflict-free addresses) and non-packed
// FP32 VALU issue rate. Prints one JSON line.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e = (x);                                                            \
    if (e != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

// Spins for `cycles` shader clocks (s_memtime counts the shader clock).
__global__ void clock_kernel(long long cycles, long long *out) {
  const long long t0 = clock64();
  long long t = t0;
  while (t - t0 < cycles) t = clock64();
  if (threadIdx.x == 0) out[blockIdx.x] = t - t0;
}

template <typename T, int U>
__global__ __launch_bounds__(256) void lds_kernel(int iters, float *out) {
  __shared__ T buf[16384 / sizeof(T) * 4];  // 64 KiB
  constexpr int n = 65536 / sizeof(T);
  for (int i = threadIdx.x; i < n; i += 256) {
    T v;
    float *f = reinterpret_cast<float *>(&v);
    for (int k = 0; k < static_cast<int>(sizeof(T) / 4); ++k) f[k] = static_cast<float>(i + k);
    buf[i] = v;
  }
  __syncthreads();
  float acc[U];
  for (int u = 0; u < U; ++u) acc[u] = 0.0f;
  int base = threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T v = buf[(base + u * 256) & (n - 1)];
      acc[u] += reinterpret_cast<const float *>(&v)[0];
    }
    base = (base + 17 * 256) & (n - 1);
  }
  float s = 0.0f;
  for (int u = 0; u < U; ++u) s += acc[u];
  if (s == 12345.678f) out[0] = s;  // keeps the loads alive
}

__global__ __launch_bounds__(256) void valu_kernel(int iters, float *out) {
  float a[16];
  for (int k = 0; k < 16; ++k) a[k] = threadIdx.x * 1e-9f + k;
  const float b = 1.000001f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = a[k] + b;
  }
  float s = 0.0f;
  for (int k = 0; k < 16; ++k) s += a[k];
  if (s == 12345.678f) out[0] = s;
}

template <typename K>
static float time_ms(K launch) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0));
  launch();
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  return ms;
}

int main() {
  hipDeviceProp_t p;
  CHECK(hipGetDeviceProperties(&p, 0));
  const int cus = p.multiProcessorCount;
  long long *d_cyc;
  float *d_out;
  CHECK(hipMalloc(&d_cyc, sizeof(long long) * cus * 4));
  CHECK(hipMalloc(&d_out, sizeof(float)));
  // shader clock: 2e9 clocks of spinning against wall time
  const long long spin = 2000000000LL;
  const float clk_ms = time_ms([&] { hipLaunchKernelGGL(clock_kernel, dim3(cus), dim3(64), 0, 0, spin, d_cyc); });
  const double ghz = spin / (clk_ms * 1e-3) / 1e9;
  const int blocks = cus * 8, iters = 4096;
  const float b32 = time_ms([&] { hipLaunchKernelGGL((lds_kernel<float, 16>), dim3(blocks), dim3(256), 0, 0, iters, d_out); });
  const float b64 = time_ms([&] { hipLaunchKernelGGL((lds_kernel<float2, 16>), dim3(blocks), dim3(256), 0, 0, iters, d_out); });
  const float b128 = time_ms([&] { hipLaunchKernelGGL((lds_kernel<float4, 8>), dim3(blocks), dim3(256), 0, 0, iters, d_out); });
  const double reads = static_cast<double>(blocks) * 256 * iters;
  const double gb32 = reads * 16 * 4 / (b32 * 1e-3) / 1e9;
  const double gb64 = reads * 16 * 8 / (b64 * 1e-3) / 1e9;
  const double gb128 = reads * 8 * 16 / (b128 * 1e-3) / 1e9;
  const int viters = 1 << 16;
  const float vms = time_ms([&] { hipLaunchKernelGGL(valu_kernel, dim3(blocks), dim3(256), 0, 0, viters, d_out); });
  const double lane_ops = static_cast<double>(blocks) * 256 * viters * 16 / (vms * 1e-3);
  std::printf("{\"device\": \"%s\", \"cus\": %d, \"clock_ghz\": %.3f, "
              "\"lds_gbs\": {\"b32\": %.0f, \"b64\": %.0f, \"b128\": %.0f}, "
              "\"lds_bytes_per_clk_per_cu\": {\"b32\": %.1f, \"b64\": %.1f, \"b128\": %.1f}, "
              "\"valu_fp32_lane_ops_per_s\": %.4g, \"valu_lane_ops_per_clk_per_cu\": %.1f, "
              "\"lds_per_block\": %zu, \"lds_per_cu\": %zu}\n",
              p.gcnArchName, cus, ghz, gb32, gb64, gb128, gb32 * 1e9 / (ghz * 1e9 * cus),
              gb64 * 1e9 / (ghz * 1e9 * cus), gb128 * 1e9 / (ghz * 1e9 * cus), lane_ops,
              lane_ops / (ghz * 1e9 * cus), static_cast<size_t>(p.sharedMemPerBlock),
              static_cast<size_t>(p.maxSharedMemoryPerMultiProcessor));
  return 0;
}
