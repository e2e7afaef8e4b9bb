// Peaks of the roofline that bench.py prices the backprojection against,
// measured on the device it runs on: the shader clock, LDS read bandwidth
// at each access width, and the issue rates of the VALU instructions the
// backprojection's inner loop is made of. Also a plain copy kernel, whose
// HBM counter readings calibrate the FETCH_SIZE / WRITE_SIZE counters.
//
// libctas_peaks.so is measurement infrastructure for bench.py; the engine
// (libctas_hip.so) does not use it.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// Spins for `cycles` shader clocks (s_memtime counts the shader clock).
__global__ void clock_kernel(long long cycles, long long *out) {
  const long long t0 = clock64();
  long long t = t0;
  while (t - t0 < cycles) t = clock64();
  if (threadIdx.x == 0) out[blockIdx.x] = t - t0;
}

template <typename T> __device__ __forceinline__ float sum_of(const T &v) {
  const float *f = reinterpret_cast<const float *>(&v);
  float a = f[0];
  for (int k = 1; k < static_cast<int>(sizeof(T) / 4); ++k) a += f[k];
  return a;
}

// U reads per thread and iteration, lane-consecutive addresses (no bank
// conflicts), each at a fixed offset from one base register (immediate
// offsets: no address arithmetic per read). An empty asm statement that
// "changes" the base register before each read keeps the compiler from
// merging two reads into one ds_read2 instruction, and a memory clobber per
// iteration from hoisting the reads out of the loop; neither emits an
// instruction. Every component of a read is summed (into four independent
// sums), so that the reads keep their full width.
template <typename T, int U>
__global__ __launch_bounds__(256) void lds_read_kernel(int iters, float *out) {
  __shared__ T buf[U * 256];
  for (int i = threadIdx.x; i < U * 256; i += 256) {
    T v;
    float *f = reinterpret_cast<float *>(&v);
    for (int k = 0; k < static_cast<int>(sizeof(T) / 4); ++k) f[k] = static_cast<float>(i + k);
    buf[i] = v;
  }
  __syncthreads();
  unsigned off = threadIdx.x * static_cast<unsigned>(sizeof(T));
  const char *base = reinterpret_cast<const char *>(buf);
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int it = 0; it < iters; ++it) {
    asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; ++u) {
      asm volatile("" : "+v"(off));
      const T v = *reinterpret_cast<const T *>(base + off + u * 256 * sizeof(T));
      acc[u & 3] += sum_of(v);
    }
  }
  const float s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  if (s == 12345.678f) out[0] = s;  // keeps the reads alive
}

// 16 independent chains of one VALU instruction kind per thread.
enum Op { OP_FADD = 0, OP_CVT = 1, OP_LSHL_ADD = 2 };

template <int OP>
__global__ __launch_bounds__(256) void valu_kernel(int iters, float *out) {
  float a[16];
  int b[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    a[k] = static_cast<float>(threadIdx.x) * 1e-3f + k;
    b[k] = static_cast<int>(threadIdx.x) + k;
  }
  const float d = 1.000001f;
  const int e = static_cast<int>(blockIdx.x) | 1;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if constexpr (OP == OP_FADD) {
        a[k] = a[k] + d;
      } else if constexpr (OP == OP_CVT) {
        b[k] = static_cast<int>(a[k]);  // v_cvt_i32_f32
        a[k] = __int_as_float(b[k] ^ e);  // v_xor (one more, counted)
      } else {
        b[k] = (b[k] << 2) + e;  // v_lshl_add_u32
      }
    }
  }
  float s = 0.0f;
  for (int k = 0; k < 16; ++k) s += a[k] + static_cast<float>(b[k]);
  if (s == 12345.678f) out[0] = s;
}

__global__ __launch_bounds__(256) void peaks_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                        long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += static_cast<long long>(gridDim.x) * 256)
    dst[i] = src[i];
}

template <typename K>
int time_ms(K launch, float *ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
  launch();
  if (hipDeviceSynchronize() != hipSuccess) return -1;
  if (hipEventRecord(e0) != hipSuccess) return -1;
  launch();
  if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(ms, e0, e1) != hipSuccess)
    return -1;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

extern "C" {

// out[12]: CUs, shader clock (GHz), LDS bytes per clock per CU for
// ds_read_b32 / b64 / b128, LDS read instructions per clock per CU for the
// same three, and VALU wave instructions per clock per CU for v_add_f32,
// v_cvt_i32_f32 (with a v_xor: the pair's rate, per instruction) and
// v_lshl_add_u32, then the LDS bytes per workgroup. Returns 0, or -1 on a
// HIP error.
int ctas_peaks_measure(double *out, int n_out) {
  if (!out || n_out < 12) return -1;
  int dev = 0, cus = 0, lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return -1;
  long long *d_cyc = nullptr;
  float *d_out = nullptr;
  if (hipMalloc(&d_cyc, sizeof(long long) * cus) != hipSuccess || hipMalloc(&d_out, sizeof(float)) != hipSuccess)
    return -1;
  int rc = 0;
  float ms = 0.0f;
  const long long spin = 1000000000LL;
  rc |= time_ms([&] { hipLaunchKernelGGL(clock_kernel, dim3(cus), dim3(64), 0, 0, spin, d_cyc); }, &ms);
  const double ghz = spin / (ms * 1e-3) / 1e9;
  const int blocks = cus * 8, iters = 4096;
  const double clk_cu = ghz * 1e9 * cus;  // CU clocks per second
  float t32 = 0, t64 = 0, t128 = 0;
  rc |= time_ms([&] { hipLaunchKernelGGL((lds_read_kernel<float, 16>), dim3(blocks), dim3(256), 0, 0, iters, d_out); }, &t32);
  rc |= time_ms([&] { hipLaunchKernelGGL((lds_read_kernel<float2, 16>), dim3(blocks), dim3(256), 0, 0, iters, d_out); }, &t64);
  rc |= time_ms([&] { hipLaunchKernelGGL((lds_read_kernel<float4, 8>), dim3(blocks), dim3(256), 0, 0, iters, d_out); }, &t128);
  const double waves = blocks * 4.0;  // wave64 per block
  const double i32 = waves * iters * 16, i64 = waves * iters * 16, i128 = waves * iters * 8;  // wave instructions
  const int viters = 1 << 14;
  float tadd = 0, tcvt = 0, tsh = 0;
  rc |= time_ms([&] { hipLaunchKernelGGL(valu_kernel<OP_FADD>, dim3(blocks), dim3(256), 0, 0, viters, d_out); }, &tadd);
  rc |= time_ms([&] { hipLaunchKernelGGL(valu_kernel<OP_CVT>, dim3(blocks), dim3(256), 0, 0, viters, d_out); }, &tcvt);
  rc |= time_ms([&] { hipLaunchKernelGGL(valu_kernel<OP_LSHL_ADD>, dim3(blocks), dim3(256), 0, 0, viters, d_out); }, &tsh);
  const double vi = waves * viters * 16.0;
  (void)hipFree(d_cyc);
  (void)hipFree(d_out);
  if (rc) return -1;
  out[0] = cus;
  out[1] = ghz;
  out[2] = i32 * 256 / (t32 * 1e-3) / clk_cu;
  out[3] = i64 * 512 / (t64 * 1e-3) / clk_cu;
  out[4] = i128 * 1024 / (t128 * 1e-3) / clk_cu;
  out[5] = i32 / (t32 * 1e-3) / clk_cu;
  out[6] = i64 / (t64 * 1e-3) / clk_cu;
  out[7] = i128 / (t128 * 1e-3) / clk_cu;
  out[8] = vi / (tadd * 1e-3) / clk_cu;
  out[9] = 2.0 * vi / (tcvt * 1e-3) / clk_cu;
  out[10] = vi / (tsh * 1e-3) / clk_cu;
  out[11] = lds;
  return 0;
}

// A plain copy of n_bytes (a multiple of 16) from src to dst on `stream`:
// the reference kernel whose counter readings scale FETCH_SIZE and
// WRITE_SIZE to bytes.
int ctas_peaks_copy(const void *src, void *dst, long long n_bytes, void *stream) {
  if (!src || !dst || n_bytes <= 0 || (n_bytes & 15)) return -1;
  const long long n = n_bytes / 16;
  hipLaunchKernelGGL(peaks_copy_kernel, dim3(4096), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float4 *>(src), static_cast<float4 *>(dst), n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
