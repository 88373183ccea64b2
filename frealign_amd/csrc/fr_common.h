// fr_common.h -- shared definitions for the gfx950 kernels and the C-ABI glue.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace fr {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(const char* fmt, ...);
#define FR_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return ::fr::fail("%s:%d %s: %s", __FILE__, __LINE__,   \
                                            #call, hipGetErrorString(e_));        \
  } while (0)
#define FR_CHECK(cond, ...)                                                       \
  do {                                                                            \
    if (!(cond)) return ::fr::fail(__VA_ARGS__);                                  \
  } while (0)

// ---------------------------------------------------------------- context
struct Ctx {
  int device = 0;
  int cus = 0;  // compute units of the device (filled on first need)
  hipStream_t stream = nullptr;
  // grow-only device scratch owned by the context
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // FFT twiddle tables, one per length (W_N^k = e^{-2 pi i k/N}, k < N)
  float2* tw[1025] = {};
  // optional per-kernel timing: HIP events recorded on the launch stream
  // around the named launches (see prof_begin / prof_end)
  bool prof_on = false;
  struct ProfRec {
    int name;
    hipEvent_t a, b;
  };
  ProfRec prof[4096];
  int nprof = 0;
  int prof_begin(int name);
  void prof_end(int slot);
  int ensure_scratch(size_t bytes);
  const float2* twiddles(int n);
  // the particle preparation's per-row-block ring sums (fr_prep.hip) for
  // calls on the context's stream: one fixed allocation for the largest
  // batch (65535 images x 16 row blocks), made on first use and kept (plans,
  // whose graphs replay on their own streams, bring their own buffer)
  float2* prep_part = nullptr;
  float2* prep_partials();
  ~Ctx();
};

// ---------------------------------------------------------------- complex
struct cf {
  float re, im;
};
__host__ __device__ inline cf cmake(float a, float b) { return cf{a, b}; }
__host__ __device__ inline cf cmul(cf a, cf b) {
  return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cf cconj(cf a) { return cf{a.re, -a.im}; }
__host__ __device__ inline cf cadd(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cf csub(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cf cscale(cf a, float s) { return cf{a.re * s, a.im * s}; }

// Wrapped frequency index -> storage row (JJ=J+1; IF (JJ.LT.1) JJ=JJ+NSAM, 0-based)
__host__ __device__ inline int wrapn(int j, int n) { return j < 0 ? j + n : j; }

// timed kernel families (fr_profile_read)
enum ProfName {
  kProfSearchGemm = 0,
  kProfInsert = 1,
  kProfCc3m = 2,
  kProfPrep = 3,
  kProfSearchList = 4,
  kProfCnn = 5,
  kProfPlan = 6,
  kProfN = 7
};

// ---------------------------------------------------------------- FFT API (fr_fft.hip)
// Batched 2D real FFTs in the reference's in-place layout: nimg images of
// NSAM x NSAM reals with row pitch NSAM+2, consecutive images `img_stride`
// floats apart.  Forward is unnormalised e^{-i}; backward is unnormalised e^{+i}
// with the row-column c2r semantic (imaginary parts of the x=0 and x=NSAM/2
// columns ignored after the column pass).
int fft_r2c_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride);
int fft_c2r_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride);
// one padded volume (NSAM+2 row pitch), in place, unnormalised
int fft_r2c_3d(Ctx& c, int nsam, float* data);
int fft_c2r_3d(Ctx& c, int nsam, float* data);
bool fft_supported(int nsam);
// the row pass alone (forward r2c of the rows, or backward c2r), for callers
// that run the column pass themselves
int fft_rows_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride, bool fwd);
// the column pass alone (sign -1 forward, +1 backward)
int fft_cols_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride, int sign);

}  // namespace fr
