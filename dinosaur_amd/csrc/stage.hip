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
//
// Ranged forms: the implicit terms may be stored only on a sub-range of the
// state (where the implicit tendency can be nonzero), and the adjoint may
// accumulate into its targets, so a hand-written reverse sweep of a whole
// IMEX step needs no separate gradient additions.
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

constexpr int kMaxTargets = 9;

// Implicit terms live on [ib, ie): element i of x[1][k] is x[1][k][i - go].
// A full-range call has ib = go = 0 and ie = n.
template <typename T>
struct StageArgs {
  const T* y0;
  const T* x[2][kMaxTerms];
  T c[2][kMaxTerms];
  int n[2];
  T dt;
  int64_t ib, ie, go;
  int64_t trl;  // > 0: Real-layout modal fields of trl = R * L elements ...
  int tl;       // ... of tl degrees per row; elements with l < m are zero
  const T* ls;  // non-null: the result is multiplied by ls[i % sl] (a
  int64_t sl;   // per-degree filter applied to the output, one rounding)
  int keep_tri;  // 1: entries below the triangle are not written
};

template <typename T>
struct TargetArgs {
  const T* g;
  T* out[kMaxTargets];
  T c[kMaxTargets];
  int kind[kMaxTargets];
  int n;
  T dt;
  int64_t ib, ie, go;
  int64_t trl;
  int tl;
};

// Elements i .. i + w - 1 all lie below the triangle (l < m, m = (r + 1) / 2
// for row r of the Real layout) of their field.  The triangle is a prefix of
// each row and m < L, so it suffices that the first element's l + w - 1 < m.
__device__ __forceinline__ bool below_triangle(int64_t i, int w, int64_t trl, int tl) {
  if (trl <= 0) return false;
  const int64_t j = i % trl;
  const int r = (int)(j / tl), l = (int)(j - (int64_t)r * tl);
  return l + w - 1 < ((r + 1) >> 1);
}

// Bit e set: element i + e (e < W) lies below the triangle of its field (the
// vector may run into the next row, or the next field's first row).
template <int W>
__device__ __forceinline__ unsigned below_mask(int64_t i, int64_t trl, int tl) {
  if (trl <= 0) return 0u;
  const int64_t j = i % trl;
  const int R = (int)(trl / tl);
  int r = (int)(j / tl), l = (int)(j - (int64_t)r * tl);
  unsigned mask = 0u;
#pragma unroll
  for (int e = 0; e < W; ++e) {
    if (l == tl) {
      l = 0;
      r = r + 1 == R ? 0 : r + 1;
    }
    if (l < ((r + 1) >> 1)) mask |= 1u << e;
    ++l;
  }
  return mask;
}

template <typename T>
struct AdjointArgs {
  const T* g;
  T* out[2][kMaxTerms];
  T c[2][kMaxTerms];
  int n[2];
  T dt;
};

// U is T (scalar) or Vec<T>::type (16-byte vector); the arithmetic is
// componentwise either way.  idx indexes U elements of the group's arrays.
template <typename T, typename U>
__device__ __forceinline__ U group_sum(const StageArgs<T>& a, int grp, int64_t idx) {
  auto ld = [](const T* p, int64_t k) { return reinterpret_cast<const U*>(p)[k]; };
  U s = a.c[grp][0] * ld(a.x[grp][0], idx);
#pragma unroll
  for (int j = 1; j < kMaxTerms; ++j)
    if (j < a.n[grp]) s = s + a.c[grp][j] * ld(a.x[grp][j], idx);
  return s;
}

template <typename T>
__device__ __forceinline__ T combine_scalar(const StageArgs<T>& a, int64_t i) {
  if (below_triangle(i, 1, a.trl, a.tl)) return (T)0;
  T y = a.y0[i];
  if (a.n[0]) y = y + a.dt * group_sum<T, T>(a, 0, i);
  if (a.n[1] && i >= a.ib && i < a.ie) y = y + a.dt * group_sum<T, T>(a, 1, i - a.go);
  if (a.ls) y = y * a.ls[i % a.sl];
  return y;
}

// VEC: every full-size pointer 16-byte aligned, implicit pointers too, and
// go a multiple of the vector width; thread k handles vector k (vectors that
// straddle a range end go component by component), and the n % W tail is done
// by scalar threads after the last full vector.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void stage_kernel(StageArgs<T> a, T* __restrict__ out,
                                                    int64_t n) {
  typedef typename Vec<T>::type V;
  constexpr int W = VEC ? Vec<T>::W : 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = n / W;
  if (k < nv) {
    if constexpr (VEC) {
      const int64_t i0 = k * W;
      const unsigned bm = below_mask<W>(i0, a.trl, a.tl);
      if (bm == (1u << W) - 1u) {  // every operand is zero there
        if (!a.keep_tri) reinterpret_cast<V*>(out)[k] = V{};
        return;
      }
      V y = reinterpret_cast<const V*>(a.y0)[k];
      if (a.n[0]) y = y + a.dt * group_sum<T, V>(a, 0, k);
      if (a.n[1]) {
        if (i0 >= a.ib && i0 + W <= a.ie) {
          y = y + a.dt * group_sum<T, V>(a, 1, (i0 - a.go) / W);
        } else if (i0 + W > a.ib && i0 < a.ie) {
#pragma unroll
          for (int e = 0; e < W; ++e) {
            const int64_t i = i0 + e;
            if (i >= a.ib && i < a.ie) y[e] = y[e] + a.dt * group_sum<T, T>(a, 1, i - a.go);
          }
        }
      }
      if (a.ls) {
        int64_t l = i0 % a.sl;
#pragma unroll
        for (int e = 0; e < W; ++e) {
          y[e] = y[e] * a.ls[l];
          l = l + 1 == a.sl ? 0 : l + 1;
        }
      }
      // a vector that the triangle cuts: its entries below it are zero too,
      // whatever the operands hold there (a stage value may leave them
      // unwritten, and the implicit terms computed from it then hold
      // anything there)
#pragma unroll
      for (int e = 0; e < W; ++e)
        if ((bm >> e) & 1u) y[e] = (T)0;
      reinterpret_cast<V*>(out)[k] = y;
    } else if (!(a.keep_tri && below_triangle(k, 1, a.trl, a.tl))) {
      out[k] = combine_scalar(a, k);
    }
  } else if (VEC && k - nv < n - nv * W) {
    const int64_t i = nv * W + (k - nv);
    if (!(a.keep_tri && below_triangle(i, 1, a.trl, a.tl))) out[i] = combine_scalar(a, i);
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

template <typename T>
__device__ __forceinline__ void target_scalar(const TargetArgs<T>& a, int64_t i) {
  const bool tri = below_triangle(i, 1, a.trl, a.tl);
  const T g = tri ? (T)0 : a.g[i];
  const T t = a.dt * g;
#pragma unroll
  for (int j = 0; j < kMaxTargets; ++j) {
    if (j >= a.n) break;
    const int kind = a.kind[j];
    const T v = (kind & DINO_TARGET_IDENTITY) ? g : a.c[j] * t;
    if (tri && (kind & DINO_TARGET_ACCUMULATE)) continue;  // adds zero
    if (kind & DINO_TARGET_RANGED) {
      if (i < a.ib || i >= a.ie) continue;
      T* o = a.out[j] + (i - a.go);
      *o = (kind & DINO_TARGET_ACCUMULATE) ? *o + v : v;
    } else {
      T* o = a.out[j] + i;
      *o = (kind & DINO_TARGET_ACCUMULATE) ? *o + v : v;
    }
  }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void targets_kernel(TargetArgs<T> a, int64_t n) {
  typedef typename Vec<T>::type V;
  constexpr int W = VEC ? Vec<T>::W : 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = n / W;
  if (k < nv) {
    if constexpr (VEC) {
      const int64_t i0 = k * W;
      const bool tri = below_triangle(i0, W, a.trl, a.tl);
      const V g = tri ? V{} : reinterpret_cast<const V*>(a.g)[k];
      const V t = a.dt * g;
      const bool inside = i0 >= a.ib && i0 + W <= a.ie;
      const bool touch = i0 + W > a.ib && i0 < a.ie;
#pragma unroll
      for (int j = 0; j < kMaxTargets; ++j) {
        if (j >= a.n) break;
        const int kind = a.kind[j];
        const bool acc = (kind & DINO_TARGET_ACCUMULATE) != 0;
        const V v = (kind & DINO_TARGET_IDENTITY) ? g : a.c[j] * t;
        if (tri && acc) continue;  // adds zero
        if (!(kind & DINO_TARGET_RANGED)) {
          V* o = reinterpret_cast<V*>(a.out[j]) + k;
          *o = acc ? *o + v : v;
        } else if (inside) {
          V* o = reinterpret_cast<V*>(a.out[j]) + (i0 - a.go) / W;
          *o = acc ? *o + v : v;
        } else if (touch) {
#pragma unroll
          for (int e = 0; e < W; ++e) {
            const int64_t i = i0 + e;
            if (i >= a.ib && i < a.ie) {
              T* o = a.out[j] + (i - a.go);
              *o = acc ? *o + v[e] : v[e];
            }
          }
        }
      }
    } else {
      target_scalar(a, k);
    }
  } else if (VEC && k - nv < n - nv * W) {
    target_scalar(a, nv * W + (k - nv));
  }
}

// Gather: one cotangent from several sources, out = v_0 + v_1 + ... in source
// order, v_k = src_k (identity) or c_k (dt src_k) -- the same roundings and
// order as a first targets pass writing v_0 followed by accumulating passes.
// Output element o is state element i = o + off (off = g_origin for a ranged
// output, else 0); a ranged output is written only where i is in [ib, ie).
constexpr int kMaxSources = 8;

template <typename T>
struct GatherArgs {
  const T* src[kMaxSources];
  T c[kMaxSources];
  int flags[kMaxSources];  // DINO_GATHER_IDENTITY | _SCALED | _IMPLICIT_RANGE
  int n;
  T dt;
  int64_t off, ib, ie;
  int64_t trl;
  int tl;
  const T* ls;  // per-degree scale of the DINO_GATHER_SCALED sources, ls[i % sl]
  int64_t sl;
  int64_t rb, re;  // the implicit range: DINO_GATHER_IMPLICIT_RANGE sources hold
                   // values there only (zero elsewhere, not read)
  int keep_tri;    // 1: entries below the triangle are not written
};

// U = T: idx is the element; U = vector: idx is the vector (elements W idx ..)
template <typename T, typename U>
__device__ __forceinline__ U gather_sum(const GatherArgs<T>& a, int64_t idx) {
  constexpr int W = sizeof(U) / sizeof(T);
  auto term = [&](int k) {
    const int64_t e0 = idx * W;
    const bool rng = (a.flags[k] & DINO_GATHER_IMPLICIT_RANGE) != 0;
    U g{};
    if (!rng || (e0 + W > a.rb && e0 < a.re)) {
      g = reinterpret_cast<const U*>(a.src[k])[idx];
      if constexpr (W > 1) {
        if (rng && !(e0 >= a.rb && e0 + W <= a.re)) {
#pragma unroll
          for (int e = 0; e < W; ++e)
            if (e0 + e < a.rb || e0 + e >= a.re) g[e] = (T)0;
        }
      }
    }
    if (a.flags[k] & DINO_GATHER_SCALED) {
      if constexpr (W == 1) {
        g = g * a.ls[idx % a.sl];
      } else {
        int64_t l = (idx * W) % a.sl;
#pragma unroll
        for (int e = 0; e < W; ++e) {
          g[e] = g[e] * a.ls[l];
          l = l + 1 == a.sl ? 0 : l + 1;
        }
      }
    }
    return (a.flags[k] & DINO_GATHER_IDENTITY) ? g : a.c[k] * (a.dt * g);
  };
  U s = term(0);
#pragma unroll
  for (int k = 1; k < kMaxSources; ++k)
    if (k < a.n) s = s + term(k);
  return s;
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs<T> a, T* __restrict__ out,
                                                     int64_t n_out) {
  typedef typename Vec<T>::type V;
  constexpr int W = VEC ? Vec<T>::W : 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = n_out / W;
  auto one = [&](int64_t o) {
    const int64_t i = o + a.off;
    if (i < a.ib || i >= a.ie) return;
    const bool tri = below_triangle(i, 1, a.trl, a.tl);
    if (tri && a.keep_tri) return;
    out[o] = tri ? (T)0 : gather_sum<T, T>(a, i);
  };
  if (k < nv) {
    if constexpr (VEC) {
      const int64_t o0 = k * W, i0 = o0 + a.off;
      if (i0 >= a.ib && i0 + W <= a.ie) {
        const bool tri = below_triangle(i0, W, a.trl, a.tl);
        if (!(tri && a.keep_tri))
          reinterpret_cast<V*>(out)[k] = tri ? V{} : gather_sum<T, V>(a, i0 / W);
      } else {
#pragma unroll
        for (int e = 0; e < W; ++e) one(o0 + e);
      }
    } else {
      one(k);
    }
  } else if (VEC && k - nv < n_out - nv * W) {
    one(nv * W + (k - nv));
  }
}

inline bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

template <typename T>
int launch_stage(const void* y0, int n_ex, const void* const* f, const double* a, int n_im,
                 const void* const* g, const double* b, double dt, void* out, int64_t n,
                 int64_t ib, int64_t ie, int64_t go, hipStream_t s, int64_t trl = 0,
                 int tl = 0, const void* ls = nullptr, int64_t sl = 0, bool keep_tri = false) {
  StageArgs<T> args{};
  args.keep_tri = keep_tri && trl > 0 ? 1 : 0;
  args.trl = trl;
  args.tl = tl;
  args.ls = static_cast<const T*>(ls);
  args.sl = sl;
  args.y0 = static_cast<const T*>(y0);
  args.n[0] = n_ex;
  args.n[1] = n_im;
  args.dt = (T)dt;
  args.ib = ib;
  args.ie = ie;
  args.go = go;
  bool vec = aligned16(y0) && aligned16(out) && go % Vec<T>::W == 0;
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

template <typename T>
int launch_targets(const void* gbar, int nt, void* const* outp, const double* coef,
                   const int* kind, double dt, int64_t n, int64_t ib, int64_t ie, int64_t go,
                   hipStream_t s, int64_t trl, int tl) {
  TargetArgs<T> args{};
  args.trl = trl;
  args.tl = tl;
  args.g = static_cast<const T*>(gbar);
  args.n = nt;
  args.dt = (T)dt;
  args.ib = ib;
  args.ie = ie;
  args.go = go;
  bool vec = aligned16(gbar) && go % Vec<T>::W == 0;
  for (int j = 0; j < nt; ++j) {
    args.out[j] = static_cast<T*>(outp[j]);
    args.c[j] = (T)coef[j];
    args.kind[j] = kind[j];
    vec = vec && aligned16(outp[j]);
  }
  const int W = vec ? Vec<T>::W : 1;
  const int64_t threads = n / W + (n % W);
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  if (vec)
    targets_kernel<T, true><<<blocks, 256, 0, s>>>(args, n);
  else
    targets_kernel<T, false><<<blocks, 256, 0, s>>>(args, n);
  DINO_LAUNCH_CHECK("imex stage targets kernel");
  return DINO_OK;
}

template <typename T>
int launch_gather(const void* const* src, const double* coef, const int* flags, int ns,
                  double dt, void* out, int64_t n_out, int64_t off, int64_t ib, int64_t ie,
                  hipStream_t s, int64_t trl, int tl, const void* ls, int64_t sl, int64_t rb,
                  int64_t re, bool keep_tri) {
  GatherArgs<T> args{};
  args.keep_tri = keep_tri && trl > 0 ? 1 : 0;
  args.ls = static_cast<const T*>(ls);
  args.sl = sl;
  args.rb = rb;
  args.re = re;
  args.n = ns;
  args.dt = (T)dt;
  args.off = off;
  args.ib = ib;
  args.ie = ie;
  args.trl = trl;
  args.tl = tl;
  bool vec = aligned16(out) && off % Vec<T>::W == 0;
  for (int k = 0; k < ns; ++k) {
    args.src[k] = static_cast<const T*>(src[k]);
    args.c[k] = (T)coef[k];
    args.flags[k] = flags[k];
    vec = vec && aligned16(src[k]);
  }
  const int W = vec ? Vec<T>::W : 1;
  const int64_t threads = n_out / W + (n_out % W);
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  if (vec)
    gather_kernel<T, true><<<blocks, 256, 0, s>>>(args, static_cast<T*>(out), n_out);
  else
    gather_kernel<T, false><<<blocks, 256, 0, s>>>(args, static_cast<T*>(out), n_out);
  DINO_LAUNCH_CHECK("imex gather kernel");
  return DINO_OK;
}

int check_range(const char* name, int64_t n, int64_t ib, int64_t ie, int64_t go) {
  DINO_REQUIRE(0 <= go && go <= ib && ib <= ie && ie <= n, DINO_EINVAL,
               "%s: need 0 <= g_origin <= im_begin <= im_end <= n (got %lld %lld %lld %lld)",
               name, (long long)go, (long long)ib, (long long)ie, (long long)n);
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
    return launch_stage<double>(y0, n_ex, f, a, n_im, g, b, dt, out, n, 0, n, 0,
                                as_stream(stream));
  return launch_stage<float>(y0, n_ex, f, a, n_im, g, b, dt, out, n, 0, n, 0,
                             as_stream(stream));
}

namespace dino {
namespace {

int stage_ranged(const char* name, const void* y0, int n_ex, const void* const* f,
                 const double* a, int n_im, const void* const* g, const double* b, double dt,
                 void* out, int64_t n, int64_t im_begin, int64_t im_end, int64_t g_origin,
                 int64_t field_size, int degrees, const void* scale, int64_t scale_len,
                 int dtype, void* stream) {
  DINO_REQUIRE(degrees == 0 || (degrees > 0 && field_size > 0 && field_size % degrees == 0 &&
                                (field_size / degrees) % 2 == 1 && n % field_size == 0),
               DINO_EINVAL, "%s: invalid Real-layout field shape (%lld, %d)", name,
               (long long)field_size, degrees);
  DINO_REQUIRE(!scale || (scale_len > 0 && n % scale_len == 0), DINO_EINVAL,
               "%s: the scale length %lld must divide the size %lld", name,
               (long long)scale_len, (long long)n);
  const bool keep_tri = (dtype & DINO_STAGE_TRIANGLE_ONLY) != 0;
  dtype &= ~DINO_STAGE_TRIANGLE_ONLY;
  int rc = check_args(name, n_ex, n_im, a, b, n, dtype);
  if (rc != DINO_OK) return rc;
  rc = check_range(name, n, im_begin, im_end, g_origin);
  if (rc != DINO_OK) return rc;
  DINO_REQUIRE(y0 && out && (n_ex == 0 || f) && (n_im == 0 || g), DINO_EINVAL,
               "%s: NULL argument", name);
  for (int j = 0; j < n_ex; ++j) DINO_REQUIRE(f[j], DINO_EINVAL, "%s: NULL f", name);
  for (int j = 0; j < n_im; ++j) DINO_REQUIRE(g[j], DINO_EINVAL, "%s: NULL g", name);
  if (n == 0) return DINO_OK;
  const int64_t trl = degrees ? field_size : 0;
  if (dtype == DINO_F64)
    return launch_stage<double>(y0, n_ex, f, a, n_im, g, b, dt, out, n, im_begin, im_end,
                                g_origin, as_stream(stream), trl, degrees, scale, scale_len,
                                keep_tri);
  return launch_stage<float>(y0, n_ex, f, a, n_im, g, b, dt, out, n, im_begin, im_end,
                             g_origin, as_stream(stream), trl, degrees, scale, scale_len,
                             keep_tri);
}

}  // namespace
}  // namespace dino

extern "C" int dino_imex_stage_ranged(const void* y0, int n_ex, const void* const* f,
                                      const double* a, int n_im, const void* const* g,
                                      const double* b, double dt, void* out, int64_t n,
                                      int64_t im_begin, int64_t im_end, int64_t g_origin,
                                      int64_t field_size, int degrees, int dtype,
                                      void* stream) {
  return stage_ranged("dino_imex_stage_ranged", y0, n_ex, f, a, n_im, g, b, dt, out, n,
                      im_begin, im_end, g_origin, field_size, degrees, nullptr, 0, dtype,
                      stream);
}

extern "C" int dino_imex_stage_ranged_scaled(const void* y0, int n_ex, const void* const* f,
                                             const double* a, int n_im, const void* const* g,
                                             const double* b, double dt, void* out, int64_t n,
                                             int64_t im_begin, int64_t im_end,
                                             int64_t g_origin, int64_t field_size, int degrees,
                                             const void* degree_scale, int64_t scale_len,
                                             int dtype, void* stream) {
  DINO_REQUIRE(degree_scale, DINO_EINVAL, "dino_imex_stage_ranged_scaled: NULL scale");
  return stage_ranged("dino_imex_stage_ranged_scaled", y0, n_ex, f, a, n_im, g, b, dt, out, n,
                      im_begin, im_end, g_origin, field_size, degrees, degree_scale, scale_len,
                      dtype, stream);
}

extern "C" int dino_imex_stage_adjoint_targets(const void* gbar, int n_targets,
                                               void* const* out, const double* coef,
                                               const int* kind, double dt, int64_t n,
                                               int64_t im_begin, int64_t im_end,
                                               int64_t g_origin, int64_t field_size,
                                               int degrees, int dtype, void* stream) {
  const char* name = "dino_imex_stage_adjoint_targets";
  DINO_REQUIRE(degrees == 0 || (degrees > 0 && field_size > 0 && field_size % degrees == 0 &&
                                (field_size / degrees) % 2 == 1 && n % field_size == 0),
               DINO_EINVAL, "%s: invalid Real-layout field shape (%lld, %d)", name,
               (long long)field_size, degrees);
  DINO_REQUIRE(n_targets >= 0 && n_targets <= kMaxTargets, DINO_EINVAL,
               "%s: at most %d targets (got %d)", name, kMaxTargets, n_targets);
  DINO_REQUIRE(gbar && (n_targets == 0 || (out && coef && kind)), DINO_EINVAL,
               "%s: NULL argument", name);
  DINO_REQUIRE(n >= 0, DINO_EINVAL, "%s: negative size", name);
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "%s: invalid dtype %d",
               name, dtype);
  int rc = check_range(name, n, im_begin, im_end, g_origin);
  if (rc != DINO_OK) return rc;
  for (int j = 0; j < n_targets; ++j) {
    DINO_REQUIRE(out[j], DINO_EINVAL, "%s: NULL target %d", name, j);
    DINO_REQUIRE((kind[j] & ~(DINO_TARGET_ACCUMULATE | DINO_TARGET_IDENTITY |
                              DINO_TARGET_RANGED)) == 0,
                 DINO_EINVAL, "%s: invalid kind %d", name, kind[j]);
  }
  if (n == 0 || n_targets == 0) return DINO_OK;
  if (dtype == DINO_F64)
    return launch_targets<double>(gbar, n_targets, out, coef, kind, dt, n, im_begin, im_end,
                                  g_origin, as_stream(stream), degrees ? field_size : 0,
                                  degrees);
  return launch_targets<float>(gbar, n_targets, out, coef, kind, dt, n, im_begin, im_end,
                               g_origin, as_stream(stream), degrees ? field_size : 0, degrees);
}

extern "C" int dino_imex_adjoint_gather(const void* const* src, const double* coef,
                                        const int* flags, int n_src, double dt, void* out,
                                        int64_t n, int ranged, int64_t im_begin,
                                        int64_t im_end, int64_t g_origin, int64_t field_size,
                                        int degrees, const void* degree_scale,
                                        int64_t scale_len, int dtype, void* stream) {
  const char* name = "dino_imex_adjoint_gather";
  const bool keep_tri = (dtype & DINO_STAGE_TRIANGLE_ONLY) != 0;
  dtype &= ~DINO_STAGE_TRIANGLE_ONLY;
  DINO_REQUIRE(degrees == 0 || (degrees > 0 && field_size > 0 && field_size % degrees == 0 &&
                                (field_size / degrees) % 2 == 1 && n % field_size == 0),
               DINO_EINVAL, "%s: invalid Real-layout field shape (%lld, %d)", name,
               (long long)field_size, degrees);
  DINO_REQUIRE(n_src >= 1 && n_src <= kMaxSources, DINO_EINVAL,
               "%s: need 1 to %d sources (got %d)", name, kMaxSources, n_src);
  DINO_REQUIRE(src && coef && flags && out, DINO_EINVAL, "%s: NULL argument", name);
  DINO_REQUIRE(n >= 0, DINO_EINVAL, "%s: negative size", name);
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "%s: invalid dtype %d",
               name, dtype);
  for (int k = 0; k < n_src; ++k) {
    DINO_REQUIRE(src[k], DINO_EINVAL, "%s: NULL source %d", name, k);
    DINO_REQUIRE((flags[k] & ~(DINO_GATHER_IDENTITY | DINO_GATHER_SCALED |
                               DINO_GATHER_IMPLICIT_RANGE)) == 0,
                 DINO_EINVAL, "%s: invalid flags %d", name, flags[k]);
    if (flags[k] & DINO_GATHER_IMPLICIT_RANGE) {
      int rc = check_range(name, n, im_begin, im_end, g_origin);
      if (rc != DINO_OK) return rc;
    }
    DINO_REQUIRE(!(flags[k] & DINO_GATHER_SCALED) ||
                     (degree_scale && scale_len > 0 && n % scale_len == 0),
                 DINO_EINVAL, "%s: a scaled source needs a scale whose length divides %lld",
                 name, (long long)n);
  }
  int64_t ib = 0, ie = n, off = 0, n_out = n;
  if (ranged) {
    int rc = check_range(name, n, im_begin, im_end, g_origin);
    if (rc != DINO_OK) return rc;
    ib = im_begin;
    ie = im_end;
    off = g_origin;
    n_out = im_end - g_origin;
  }
  if (n_out == 0) return DINO_OK;
  const int64_t trl = degrees ? field_size : 0;
  if (dtype == DINO_F64)
    return launch_gather<double>(src, coef, flags, n_src, dt, out, n_out, off, ib, ie,
                                 as_stream(stream), trl, degrees, degree_scale, scale_len,
                                 im_begin, im_end, keep_tri);
  return launch_gather<float>(src, coef, flags, n_src, dt, out, n_out, off, ib, ie,
                              as_stream(stream), trl, degrees, degree_scale, scale_len,
                              im_begin, im_end, keep_tri);
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
