// Vertical regridding of nodal columns between hybrid and sigma levels, the
// step that brings reanalysis fields (e.g. 137 hybrid levels) onto the
// model's sigma levels before digital-filter initialisation.  Replaces the
// per-column vmaps of dinosaur/vertical_interpolation.py:
//   regrid_hybrid_to_sigma   :366-401  (conservative, in sigma = a / ps + b)
//   regrid_hybrid_to_hybrid  :404-440  (conservative, in pressure a + b ps)
//   interp_hybrid_to_sigma   :303-326  (linear between layer centres, one
//                                       linearly extrapolated cell at each end)
// which build a dense (target x source) weight matrix per column.  Here one
// thread owns one (field, column): it forms the column's level boundaries
// from (a, b) and the column's surface pressure and walks source and target
// levels together (both are monotonic), so the work per column is
// O(K_source + K_target) and each source value is read once or twice.
// Loads and stores are coalesced along columns (fields are (levels, columns)
// with columns contiguous).  Boundaries and weights are formed in FP64.
#include <cmath>
#include <cstring>
#include <vector>

#include "dino_common.h"

struct dino_vregrid {
  int mode, Ks, Kt;
  double* c64;  // sa[Ks+1], sb[Ks+1], ta[Kt+1], tb[Kt+1]
};

namespace dino {
namespace {

struct VParams {
  int mode, Ks, Kt;
  int64_t cols, per_ps;  // fields sharing one surface-pressure slice
};

template <typename T>
__global__ __launch_bounds__(256) void vregrid_kernel(const T* __restrict__ in,
                                                      const T* __restrict__ ps_all,
                                                      T* __restrict__ out,
                                                      const double* __restrict__ c, VParams p) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t f = blockIdx.y;
  if (col >= p.cols) return;
  const int Ks = p.Ks, Kt = p.Kt;
  const double* sa = c;
  const double* sb = c + Ks + 1;
  const double* ta = sb + Ks + 1;
  const double* tb = ta + Kt + 1;
  const double ps = (double)ps_all[(f / p.per_ps) * p.cols + col];
  const T* x = in + f * Ks * p.cols + col;
  T* y = out + f * Kt * p.cols + col;
  if (p.mode == DINO_VREGRID_LINEAR_SIGMA) {
    // source centres in sigma, extended by one cell at each end; target
    // centres as the caller's sigma coordinates hold them (ta)
    auto sc = [&](int k) { return 0.5 * ((sa[k] + sa[k + 1]) / ps + sb[k] + sb[k + 1]); };
    // extended abscissa / ordinate index e in [0, Ks + 1]: e = 0 and
    // e = Ks + 1 are the extrapolated cells
    auto xe = [&](int e) -> double {
      if (e == 0) return Ks > 1 ? 2 * sc(0) - sc(1) : sc(0);
      if (e == Ks + 1) return Ks > 1 ? 2 * sc(Ks - 1) - sc(Ks - 2) : sc(Ks - 1);
      return sc(e - 1);
    };
    auto fe = [&](int e) -> double {
      if (e == 0) return Ks > 1 ? 2.0 * (double)x[0] - (double)x[p.cols] : (double)x[0];
      if (e == Ks + 1)
        return Ks > 1 ? 2.0 * (double)x[(int64_t)(Ks - 1) * p.cols] -
                            (double)x[(int64_t)(Ks - 2) * p.cols]
                      : (double)x[(int64_t)(Ks - 1) * p.cols];
      return (double)x[(int64_t)(e - 1) * p.cols];
    };
    int e = 0;
    for (int t = 0; t < Kt; ++t) {
      const double xt = ta[t];
      double v;
      if (xt < xe(0) || xt > xe(Ks + 1)) {
        v = NAN;
      } else {
        while (e < Ks && xe(e + 1) < xt) ++e;
        const double x0 = xe(e), x1 = xe(e + 1);
        const double w = x1 > x0 ? (xt - x0) / (x1 - x0) : 0.0;
        v = fe(e) + w * (fe(e + 1) - fe(e));
      }
      y[(int64_t)t * p.cols] = (T)v;
    }
    return;
  }
  const bool pressure = p.mode == DINO_VREGRID_CONSERVATIVE_PRESSURE;
  auto sbnd = [&](int k) { return pressure ? sa[k] + sb[k] * ps : sa[k] / ps + sb[k]; };
  auto tbnd = [&](int k) { return pressure ? ta[k] + tb[k] * ps : ta[k] / ps + tb[k]; };
  int s = 0;
  double t_lo = tbnd(0);
  for (int t = 0; t < Kt; ++t) {
    const double t_hi = tbnd(t + 1);
    while (s < Ks && sbnd(s + 1) <= t_lo) ++s;
    double acc = 0, wsum = 0;
    double s_lo = s < Ks ? sbnd(s) : 0.0;
    for (int k = s; k < Ks && s_lo < t_hi; ++k) {
      const double s_hi = sbnd(k + 1);
      const double w = fmin(t_hi, s_hi) - fmax(t_lo, s_lo);
      if (w > 0) {
        acc += w * (double)x[(int64_t)k * p.cols];
        wsum += w;
      }
      s_lo = s_hi;
    }
    // cells with no overlap are zero (the reference's nan_to_num)
    y[(int64_t)t * p.cols] = (T)(wsum > 0 ? acc / wsum : 0.0);
    t_lo = t_hi;
  }
}

}  // namespace
}  // namespace dino

using namespace dino;

extern "C" int dino_vregrid_create(int mode, int source_levels, const double* source_a,
                                   const double* source_b, int target_levels,
                                   const double* target_a, const double* target_b,
                                   dino_vregrid** out) {
  DINO_REQUIRE(out, DINO_EINVAL, "dino_vregrid_create: out is NULL");
  *out = nullptr;
  DINO_REQUIRE(mode == DINO_VREGRID_CONSERVATIVE_SIGMA ||
                   mode == DINO_VREGRID_CONSERVATIVE_PRESSURE ||
                   mode == DINO_VREGRID_LINEAR_SIGMA,
               DINO_EINVAL, "unknown regrid mode %d", mode);
  DINO_REQUIRE(source_levels >= 1 && target_levels >= 1, DINO_EINVAL, "levels must be >= 1");
  DINO_REQUIRE(source_a && source_b && target_a && target_b, DINO_EINVAL,
               "NULL boundary array");
  const int Ks = source_levels, Kt = target_levels;
  std::vector<double> h;
  h.insert(h.end(), source_a, source_a + Ks + 1);
  h.insert(h.end(), source_b, source_b + Ks + 1);
  h.insert(h.end(), target_a, target_a + Kt + 1);
  h.insert(h.end(), target_b, target_b + Kt + 1);
  // Hybrid targets can be checked only column by column (their order depends
  // on the surface pressure); sigma targets are fixed.
  if (mode == DINO_VREGRID_CONSERVATIVE_SIGMA)
    for (int k = 0; k < Kt; ++k)
      DINO_REQUIRE(target_b[k + 1] >= target_b[k], DINO_EINVAL,
                   "target sigma boundaries must increase downwards");
  if (mode == DINO_VREGRID_LINEAR_SIGMA)
    for (int k = 0; k + 1 < Kt; ++k)
      DINO_REQUIRE(target_a[k + 1] >= target_a[k], DINO_EINVAL,
                   "target sigma centres must increase downwards");
  dino_vregrid* r = new dino_vregrid();
  std::memset(r, 0, sizeof(*r));
  r->mode = mode;
  r->Ks = Ks;
  r->Kt = Kt;
  hipError_t e = hipMalloc(&r->c64, sizeof(double) * h.size());
  if (e == hipSuccess)
    e = hipMemcpy(r->c64, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (r->c64) (void)hipFree(r->c64);
    delete r;
    return hip_fail(e, "dino_vregrid_create");
  }
  *out = r;
  return DINO_OK;
}

extern "C" int dino_vregrid_destroy(dino_vregrid* r) {
  if (r) {
    if (r->c64) (void)hipFree(r->c64);
    delete r;
  }
  return DINO_OK;
}

extern "C" int dino_vregrid_apply(const dino_vregrid* r, const void* fields,
                                  const void* surface_pressure, void* out, int64_t n_fields,
                                  int64_t n_surface, int64_t columns, int dtype, void* stream) {
  DINO_REQUIRE(r, DINO_EINVAL, "dino_vregrid_apply: NULL regridder");
  DINO_REQUIRE(n_fields >= 0 && columns >= 0, DINO_EINVAL, "negative size");
  DINO_REQUIRE(n_fields <= 65535, DINO_EINVAL, "at most 65535 fields per call, got %lld",
               (long long)n_fields);
  DINO_REQUIRE(n_surface >= 1 && n_fields % n_surface == 0, DINO_EINVAL,
               "the %lld fields must split evenly over %lld surface-pressure slices",
               (long long)n_fields, (long long)n_surface);
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "unknown dtype %d", dtype);
  if (n_fields == 0 || columns == 0) return DINO_OK;
  DINO_REQUIRE(fields && surface_pressure && out, DINO_EINVAL, "NULL buffer");
  VParams p;
  p.mode = r->mode;
  p.Ks = r->Ks;
  p.Kt = r->Kt;
  p.cols = columns;
  p.per_ps = n_fields / n_surface;
  dim3 grid((unsigned)((columns + 255) / 256), (unsigned)n_fields);
  hipStream_t s = as_stream(stream);
  if (dtype == DINO_F64)
    vregrid_kernel<double><<<grid, 256, 0, s>>>(static_cast<const double*>(fields),
                                                static_cast<const double*>(surface_pressure),
                                                static_cast<double*>(out), r->c64, p);
  else
    vregrid_kernel<float><<<grid, 256, 0, s>>>(static_cast<const float*>(fields),
                                               static_cast<const float*>(surface_pressure),
                                               static_cast<float*>(out), r->c64, p);
  DINO_LAUNCH_CHECK("vertical regrid kernel");
  return DINO_OK;
}
