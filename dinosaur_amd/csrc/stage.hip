// IMEX Runge-Kutta stage combination on flat state buffers.
//
// The reference forms every stage value and the step result as
//   y = y0 + dt * (sum_j a_j f_j) + dt * (sum_k b_k g_k)
// with tree_map arithmetic (time_integration.py:351-370): each product,
// partial sum and scaling is a separate elementwise pass over the state.
// One kernel here reads every operand once and writes the result once,
// performing the same roundings in the same order per element:
//   s = fl(a_0 f_0); s = fl(s + fl(a_1 f_1)); ...; y = fl(y0 + fl(dt s));
// then the same for the implicit group.  Coefficients are rounded to the
// data type first, as a scalar operand of an elementwise product is.
//
// The adjoint reads the output cotangent once and writes the cotangent of
// every f_j and g_k: fl(a_j fl(dt gbar)).
#include "dino_common.h"

// separate roundings of every product and sum, as the elementwise passes do
#pragma clang fp contract(off)

namespace dino {
namespace {

constexpr int kMaxTerms = 4;

// 16-byte vectors: 4 floats or 2 doubles
template <typename T>
struct Vec;
template <>
struct Vec<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int W = 4;
};
template <>
struct Vec<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int W = 2;
};

template <typename T>
struct StageArgs {
  const T* y0;
  const T* x[2][kMaxTerms];
  T c[2][kMaxTerms];
  int n[2];
  T dt;
};

template <typename T>
struct AdjointArgs {
  const T* g;
  T* out[2][kMaxTerms];
  T c[2][kMaxTerms];
  int n[2];
  T dt;
};

// U is T (scalar tail) or Vec<T>::type (16-byte body); the arithmetic is
// componentwise either way.
template <typename T, typename U>
__device__ __forceinline__ U combine_one(const StageArgs<T>& a, int64_t i) {
  auto ld = [](const T* p, int64_t k) { return reinterpret_cast<const U*>(p)[k]; };
  U y = ld(a.y0, i);
#pragma unroll
  for (int grp = 0; grp < 2; ++grp) {
    if (a.n[grp] == 0) continue;
    U s = a.c[grp][0] * ld(a.x[grp][0], i);
#pragma unroll
    for (int j = 1; j < kMaxTerms; ++j)
      if (j < a.n[grp]) s = s + a.c[grp][j] * ld(a.x[grp][j], i);
    y = y + a.dt * s;
  }
  return y;
}

// VEC: every pointer 16-byte aligned; thread k handles vector k, and the
// n % W tail is done by scalar threads after the last full vector.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void stage_kernel(StageArgs<T> a, T* __restrict__ out,
                                                    int64_t n) {
  typedef typename Vec<T>::type V;
  constexpr int W = VEC ? Vec<T>::W : 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = n / W;
  if (k < nv) {
    if constexpr (VEC) reinterpret_cast<V*>(out)[k] = combine_one<T, V>(a, k);
    else out[k] = combine_one<T, T>(a, k);
  } else if (VEC && k - nv < n - nv * W) {
    const int64_t i = nv * W + (k - nv);
    out[i] = combine_one<T, T>(a, i);
  }
}

template <typename T, typename U>
__device__ __forceinline__ void adjoint_one(const AdjointArgs<T>& a, int64_t i) {
  const U t = a.dt * reinterpret_cast<const U*>(a.g)[i];
#pragma unroll
  for (int grp = 0; grp < 2; ++grp)
#pragma unroll
    for (int j = 0; j < kMaxTerms; ++j)
      if (j < a.n[grp]) reinterpret_cast<U*>(a.out[grp][j])[i] = a.c[grp][j] * t;
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void stage_adjoint_kernel(AdjointArgs<T> a, int64_t n) {
  typedef typename Vec<T>::type V;
  constexpr int W = VEC ? Vec<T>::W : 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = n / W;
  if (k < nv) {
    if constexpr (VEC) adjoint_one<T, V>(a, k);
    else adjoint_one<T, T>(a, k);
  } else if (VEC && k - nv < n - nv * W) {
    adjoint_one<T, T>(a, nv * W + (k - nv));
  }
}

inline bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

template <typename T>
int launch_stage(const void* y0, int n_ex, const void* const* f, const double* a, int n_im,
                 const void* const* g, const double* b, double dt, void* out, int64_t n,
                 hipStream_t s) {
  StageArgs<T> args{};
  args.y0 = static_cast<const T*>(y0);
  args.n[0] = n_ex;
  args.n[1] = n_im;
  args.dt = (T)dt;
  bool vec = aligned16(y0) && aligned16(out);
  for (int j = 0; j < n_ex; ++j) {
    args.x[0][j] = static_cast<const T*>(f[j]);
    args.c[0][j] = (T)a[j];
    vec = vec && aligned16(f[j]);
  }
  for (int j = 0; j < n_im; ++j) {
    args.x[1][j] = static_cast<const T*>(g[j]);
    args.c[1][j] = (T)b[j];
    vec = vec && aligned16(g[j]);
  }
  const int W = vec ? Vec<T>::W : 1;
  const int64_t threads = n / W + (n % W);
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  if (vec)
    stage_kernel<T, true><<<blocks, 256, 0, s>>>(args, static_cast<T*>(out), n);
  else
    stage_kernel<T, false><<<blocks, 256, 0, s>>>(args, static_cast<T*>(out), n);
  DINO_LAUNCH_CHECK("imex stage kernel");
  return DINO_OK;
}

template <typename T>
int launch_adjoint(const void* gout, int n_ex, const double* a, int n_im, const double* b,
                   double dt, void* const* f_bar, void* const* g_bar, int64_t n,
                   hipStream_t s) {
  AdjointArgs<T> args{};
  args.g = static_cast<const T*>(gout);
  args.n[0] = n_ex;
  args.n[1] = n_im;
  args.dt = (T)dt;
  bool vec = aligned16(gout);
  for (int j = 0; j < n_ex; ++j) {
    args.out[0][j] = static_cast<T*>(f_bar[j]);
    args.c[0][j] = (T)a[j];
    vec = vec && aligned16(f_bar[j]);
  }
  for (int j = 0; j < n_im; ++j) {
    args.out[1][j] = static_cast<T*>(g_bar[j]);
    args.c[1][j] = (T)b[j];
    vec = vec && aligned16(g_bar[j]);
  }
  const int W = vec ? Vec<T>::W : 1;
  const int64_t threads = n / W + (n % W);
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  if (vec)
    stage_adjoint_kernel<T, true><<<blocks, 256, 0, s>>>(args, n);
  else
    stage_adjoint_kernel<T, false><<<blocks, 256, 0, s>>>(args, n);
  DINO_LAUNCH_CHECK("imex stage adjoint kernel");
  return DINO_OK;
}

int check_args(const char* name, int n_ex, int n_im, const double* a, const double* b,
               int64_t n, int dtype) {
  DINO_REQUIRE(n_ex >= 0 && n_ex <= kMaxTerms && n_im >= 0 && n_im <= kMaxTerms, DINO_EINVAL,
               "%s: at most %d terms per group (got %d, %d)", name, kMaxTerms, n_ex, n_im);
  DINO_REQUIRE((n_ex == 0 || a) && (n_im == 0 || b), DINO_EINVAL, "%s: NULL coefficients",
               name);
  DINO_REQUIRE(n >= 0, DINO_EINVAL, "%s: negative size", name);
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "%s: invalid dtype %d",
               name, dtype);
  return DINO_OK;
}

}  // namespace
}  // namespace dino

using namespace dino;

extern "C" int dino_imex_stage(const void* y0, int n_ex, const void* const* f, const double* a,
                               int n_im, const void* const* g, const double* b, double dt,
                               void* out, int64_t n, int dtype, void* stream) {
  int rc = check_args("dino_imex_stage", n_ex, n_im, a, b, n, dtype);
  if (rc != DINO_OK) return rc;
  DINO_REQUIRE(y0 && out && (n_ex == 0 || f) && (n_im == 0 || g), DINO_EINVAL,
               "dino_imex_stage: NULL argument");
  for (int j = 0; j < n_ex; ++j) DINO_REQUIRE(f[j], DINO_EINVAL, "dino_imex_stage: NULL f");
  for (int j = 0; j < n_im; ++j) DINO_REQUIRE(g[j], DINO_EINVAL, "dino_imex_stage: NULL g");
  if (n == 0) return DINO_OK;
  if (dtype == DINO_F64)
    return launch_stage<double>(y0, n_ex, f, a, n_im, g, b, dt, out, n, as_stream(stream));
  return launch_stage<float>(y0, n_ex, f, a, n_im, g, b, dt, out, n, as_stream(stream));
}

extern "C" int dino_imex_stage_adjoint(const void* out_cot, int n_ex, const double* a,
                                       int n_im, const double* b, double dt, void* const* f_cot,
                                       void* const* g_cot, int64_t n, int dtype, void* stream) {
  int rc = check_args("dino_imex_stage_adjoint", n_ex, n_im, a, b, n, dtype);
  if (rc != DINO_OK) return rc;
  DINO_REQUIRE(out_cot && (n_ex == 0 || f_cot) && (n_im == 0 || g_cot), DINO_EINVAL,
               "dino_imex_stage_adjoint: NULL argument");
  for (int j = 0; j < n_ex; ++j)
    DINO_REQUIRE(f_cot[j], DINO_EINVAL, "dino_imex_stage_adjoint: NULL f cotangent");
  for (int j = 0; j < n_im; ++j)
    DINO_REQUIRE(g_cot[j], DINO_EINVAL, "dino_imex_stage_adjoint: NULL g cotangent");
  if (n == 0) return DINO_OK;
  if (dtype == DINO_F64)
    return launch_adjoint<double>(out_cot, n_ex, a, n_im, b, dt, f_cot, g_cot, n,
                                  as_stream(stream));
  return launch_adjoint<float>(out_cot, n_ex, a, n_im, b, dt, f_cot, g_cot, n,
                               as_stream(stream));
}
