// fr_stats.hip -- volume normalisation and FP64 per-shell statistics.
//
// launch_shell_sums: the FSC/PSSNR shell sums of SHELTEST (sheltest.f:48-88),
//   accumulated in FP64 on the GPU (per-block LDS accumulators, then one FP64
//   atomic per shell and block).
// launch_storeshift: the half-map normalisation of STORESHIFT
//   (storeshift.f:43-66, IFSC=0, NNSTAT=0), computed in FP64 as the reference
//   does (COMPLEX / DOUBLE promotes to double precision).
//
// The no-allocation form of the shell sums keeps its accumulators in
// `scratch`, a file-scope device array of kNoAllocShells x 6 doubles (24,960
// bytes) that every device the library runs on holds until the process ends.
// Its device address is taken once per device, at context creation, with
// that device current; kernels and copies use that address only.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fr_common.h"
#include "fr_kernels.h"

namespace fr {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxShell = kNoAllocShells;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[6 * kNoAllocShells];

std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

__global__ void k_zero_f64(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__device__ inline int isqrt_floor(int v) {
  int r = (int)sqrtf((float)v);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

__global__ void __launch_bounds__(kThreads)
    k_shell_sums(int nsam, int istp, int nstp, const float2* __restrict__ a,
                 const float2* __restrict__ d, const float* __restrict__ s,
                 const float* __restrict__ v, double* acc) {
  __shared__ double sh[kMaxShell * 6];
  for (int k = threadIdx.x; k < nstp * 6; k += blockDim.x) sh[k] = 0.0;
  __syncthreads();
  const int jc = nsam / 2 + 1;
  const size_t nvox = (size_t)jc * nsam * nsam;
  for (size_t id = blockIdx.x * (size_t)kThreads + threadIdx.x; id < nvox;
       id += (size_t)gridDim.x * kThreads) {
    int l = (int)(id % jc);
    size_t r = id / jc;
    int m = (int)(r % nsam), n = (int)(r / nsam);
    int mm = m >= jc ? m - nsam : m, nn = n >= jc ? n - nsam : n;
    int frad2 = l * l + mm * mm + nn * nn;
    if (frad2 >= jc * jc) continue;
    int shell = isqrt_floor(frad2) / istp;
    if (shell >= nstp) continue;
    float2 va = a[id], vd = d[id];
    float a1 = hypotf(va.x, va.y), a2 = hypotf(vd.x, vd.y);
    double* acc_s = sh + 6 * shell;
    atomicAdd(&acc_s[5], 1.0);  // NPIXT
    if (a1 * a2 != 0.0f) {
      float re = va.x * vd.x + va.y * vd.y;  // REAL(A*CONJG(D))
      atomicAdd(&acc_s[0], (double)re);
      atomicAdd(&acc_s[1], (double)a1 * (double)a1);
      atomicAdd(&acc_s[2], (double)a2 * (double)a2);
      atomicAdd(&acc_s[3], (double)s[id] + (double)v[id]);
      atomicAdd(&acc_s[4], 1.0);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nstp * 6; k += blockDim.x)
    if (sh[k] != 0.0) unsafeAtomicAdd(&acc[k], sh[k]);
}

__global__ void __launch_bounds__(kThreads)
    k_storeshift(int nsam, double std_, double vtd, int iredun, const float2* __restrict__ a,
                 const float* __restrict__ s, const float2* __restrict__ d,
                 const float* __restrict__ v, float2* b, float2* c) {
  const int jc = nsam / 2 + 1;
  const size_t nvox = (size_t)jc * nsam * nsam;
  const double nsam3 = (double)nsam * nsam * nsam;
  for (size_t id = blockIdx.x * (size_t)kThreads + threadIdx.x; id < nvox;
       id += (size_t)gridDim.x * kThreads) {
    int l = (int)(id % jc);
    size_t r = id / jc;
    int m = (int)(r % nsam), n = (int)(r / nsam);
    int mm = m >= jc ? m - nsam : m, nn = n >= jc ? n - nsam : n;
    float sg = ((l + mm + nn) & 1) ? -1.0f : 1.0f;
    double db = (double)s[id] + std_ * iredun;
    double dv = (double)v[id] + vtd * iredun;
    float2 va = a[id], vd = d[id];
    b[id] = make_float2((float)((double)(va.x * sg) / db / nsam3),
                        (float)((double)(va.y * sg) / db / nsam3));
    c[id] = make_float2((float)((double)(vd.x * sg) / dv / nsam3),
                        (float)((double)(vd.y * sg) / dv / nsam3));
  }
}

}  // namespace

int noalloc_scratch_init(int device) {
  FR_CHECK(device >= 0 && device < kMaxDevices, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  FR_CHECK(e == hipSuccess && p != nullptr, "hipGetSymbolAddress(scratch) failed on device %d: %s",
           device, hipGetErrorString(e));
  g_scratch_dev[device] = static_cast<double*>(p);
  return 0;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

int launch_shell_sums(Ctx& c, int nsam, int istp, int nstp, const float* a, const float* d,
                      const float* s, const float* v, double* acc, bool fill_kernel) {
  FR_CHECK(nstp <= kMaxShell, "too many shells (%d)", nstp);
  if (fill_kernel)
    hipLaunchKernelGGL(k_zero_f64, dim3((6 * nstp + 255) / 256), dim3(256), 0, c.stream, acc,
                       6 * nstp);
  else
    FR_HIP(hipMemsetAsync(acc, 0, sizeof(double) * 6 * nstp, c.stream));
  const size_t nvox = (size_t)(nsam / 2 + 1) * nsam * nsam;
  int blocks = (int)((nvox + kThreads - 1) / kThreads);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_shell_sums, dim3(blocks), dim3(kThreads), 0, c.stream, nsam, istp, nstp,
                     reinterpret_cast<const float2*>(a), reinterpret_cast<const float2*>(d), s, v,
                     acc);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_storeshift(Ctx& c, int nsam, double std_, double vtd, int iredun, const float* a,
                      const float* s, const float* d, const float* v, float* b, float* cc) {
  const size_t nvox = (size_t)(nsam / 2 + 1) * nsam * nsam;
  int blocks = (int)((nvox + kThreads - 1) / kThreads);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_storeshift, dim3(blocks), dim3(kThreads), 0, c.stream, nsam, std_, vtd,
                     iredun, reinterpret_cast<const float2*>(a), s,
                     reinterpret_cast<const float2*>(d), v, reinterpret_cast<float2*>(b),
                     reinterpret_cast<float2*>(cc));
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr
