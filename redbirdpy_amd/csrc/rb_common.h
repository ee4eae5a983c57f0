// Shared device/host helpers for the rbhip library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#define RB_OK 0
#define RB_ERR_ARG -1
#define RB_ERR_HIP -2
#define RB_ERR_SIZE -3
#define RB_ERR_PATTERN -4
#define RB_ERR_BREAKDOWN -5

void rb_set_error(const char* fmt, ...);

#define RB_HIP(expr)                                                            \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      rb_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,                 \
                   hipGetErrorString(_e));                                      \
      return RB_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

#define RB_CHECK_LAUNCH()                                                       \
  do {                                                                          \
    hipError_t _e = hipGetLastError();                                          \
    if (_e != hipSuccess) {                                                     \
      rb_set_error("%s:%d launch failed: %s", __FILE__, __LINE__,               \
                   hipGetErrorString(_e));                                      \
      return RB_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

// ---------------------------------------------------------------------------
// Minimal complex arithmetic (interleaved re, im), no library dependency.
// ---------------------------------------------------------------------------
template <typename R>
struct cplx {
  R x, y;
};

template <typename R>
__host__ __device__ inline cplx<R> mk(R a, R b) {
  cplx<R> c;
  c.x = a;
  c.y = b;
  return c;
}

template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__host__ __device__ inline cplx<R> operator*(R s, cplx<R> b) { return mk<R>(s * b.x, s * b.y); }

// Scalar traits: real type, zero, conversions, unconjugated product and |.|^2.
template <typename T>
struct sc;

template <>
struct sc<float> {
  typedef float real;
  typedef double acc;  // dot products are accumulated in double
  static constexpr bool is_complex = false;
  __host__ __device__ static float zero() { return 0.f; }
  __host__ __device__ static double to_acc(float v) { return (double)v; }
  __host__ __device__ static float from_acc(double v) { return (float)v; }
  __host__ __device__ static double norm2(float v) { return (double)v * (double)v; }
};
template <>
struct sc<double> {
  typedef double real;
  typedef double acc;
  static constexpr bool is_complex = false;
  __host__ __device__ static double zero() { return 0.0; }
  __host__ __device__ static double to_acc(double v) { return v; }
  __host__ __device__ static double from_acc(double v) { return v; }
  __host__ __device__ static double norm2(double v) { return v * v; }
};
template <>
struct sc<cplx<float>> {
  typedef float real;
  typedef cplx<double> acc;
  static constexpr bool is_complex = true;
  __host__ __device__ static cplx<float> zero() { return mk<float>(0.f, 0.f); }
  __host__ __device__ static cplx<double> to_acc(cplx<float> v) { return mk<double>(v.x, v.y); }
  __host__ __device__ static cplx<float> from_acc(cplx<double> v) { return mk<float>((float)v.x, (float)v.y); }
  __host__ __device__ static double norm2(cplx<float> v) {
    return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
};
template <>
struct sc<cplx<double>> {
  typedef double real;
  typedef cplx<double> acc;
  static constexpr bool is_complex = true;
  __host__ __device__ static cplx<double> zero() { return mk<double>(0.0, 0.0); }
  __host__ __device__ static cplx<double> to_acc(cplx<double> v) { return v; }
  __host__ __device__ static cplx<double> from_acc(cplx<double> v) { return v; }
  __host__ __device__ static double norm2(cplx<double> v) { return v.x * v.x + v.y * v.y; }
};

__host__ __device__ inline double acc_zero(double) { return 0.0; }
__host__ __device__ inline cplx<double> acc_zero(cplx<double>) { return mk<double>(0.0, 0.0); }

// multiply-accumulate helpers in the accumulation type
__device__ inline double amul(float a, float b) { return (double)a * (double)b; }
__device__ inline double amul(double a, double b) { return a * b; }
__device__ inline cplx<double> amul(cplx<float> a, cplx<float> b) {
  return mk<double>((double)a.x * b.x - (double)a.y * b.y, (double)a.x * b.y + (double)a.y * b.x);
}
__device__ inline cplx<double> amul(cplx<double> a, cplx<double> b) { return a * b; }

// x*s with s in the accumulation type, rounded to the storage type
__device__ inline float scale_to(float v, double s) { return (float)((double)v * s); }
__device__ inline double scale_to(double v, double s) { return v * s; }
__device__ inline cplx<float> scale_to(cplx<float> v, cplx<double> s) {
  cplx<double> r = mk<double>(v.x, v.y) * s;
  return mk<float>((float)r.x, (float)r.y);
}
__device__ inline cplx<double> scale_to(cplx<double> v, cplx<double> s) { return v * s; }

// coherent (agent-scope) loads of values written by other workgroups
__device__ inline double coherent_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline cplx<double> coherent_load(const cplx<double>* p) {
  const double* d = reinterpret_cast<const double*>(p);
  return mk<double>(coherent_load(d), coherent_load(d + 1));
}
__device__ inline int coherent_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline double acc_div_real(double a, double b) { return a / b; }
__device__ inline cplx<double> acc_div(cplx<double> a, cplx<double> b) {
  double d = b.x * b.x + b.y * b.y;
  return mk<double>((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ inline double acc_div(double a, double b) { return a / b; }
__device__ inline double acc_abs2(double a) { return a * a; }
__device__ inline double acc_abs2(cplx<double> a) { return a.x * a.x + a.y * a.y; }
__device__ inline double acc_neg(double a) { return -a; }
__device__ inline cplx<double> acc_neg(cplx<double> a) { return mk<double>(-a.x, -a.y); }
__device__ inline double acc_add(double a, double b) { return a + b; }
__device__ inline cplx<double> acc_add(cplx<double> a, cplx<double> b) { return a + b; }

__host__ __device__ inline double acc_from_real(double v, double) { return v; }
__host__ __device__ inline cplx<double> acc_from_real(double v, cplx<double>) { return mk<double>(v, 0.0); }
__host__ __device__ inline double acc_re(double v) { return v; }
__host__ __device__ inline double acc_re(cplx<double> v) { return v.x; }

__host__ __device__ inline float one_of(float) { return 1.f; }
__host__ __device__ inline double one_of(double) { return 1.0; }
__host__ __device__ inline cplx<float> one_of(cplx<float>) { return mk<float>(1.f, 0.f); }
__host__ __device__ inline cplx<double> one_of(cplx<double>) { return mk<double>(1.0, 0.0); }

__host__ __device__ inline float tdiv(float a, float b) { return a / b; }
__host__ __device__ inline double tdiv(double a, double b) { return a / b; }
template <typename R>
__host__ __device__ inline cplx<R> tdiv(cplx<R> a, cplx<R> b) {
  R d = b.x * b.x + b.y * b.y;
  return mk<R>((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

__host__ __device__ inline float narrow(double v, float) { return (float)v; }
__host__ __device__ inline cplx<float> narrow(cplx<double> v, cplx<float>) { return mk<float>((float)v.x, (float)v.y); }
__host__ __device__ inline double narrow(double v, double) { return v; }
__host__ __device__ inline cplx<double> narrow(cplx<double> v, cplx<double>) { return v; }
__host__ __device__ inline double widen(float v, double) { return (double)v; }
__host__ __device__ inline cplx<double> widen(cplx<float> v, cplx<double>) { return mk<double>(v.x, v.y); }
__host__ __device__ inline double widen(double v, double) { return v; }
__host__ __device__ inline cplx<double> widen(cplx<double> v, cplx<double>) { return v; }

static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Exclusive scan of int32 counts: out[0..n] with out[n] = total.
int rb_scan_i32(const int32_t* in, int32_t* out, int64_t n, hipStream_t s);
// Fill helpers (kernels, never hipMemset, so that the calls are safe in capture)
int rb_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
int rb_fill_bytes0(void* p, size_t bytes, hipStream_t s);
