// Fused nodal nonlinear terms of PrimitiveEquationsHybrid.explicit_terms on
// hybrid sigma-pressure levels p = a + b ps.
//
// One thread owns one (member, longitude, latitude) column.  Unlike sigma
// levels, the layer thickness dp_k = da_k + db_k ps, the Simmons-Burridge
// coefficients alpha_k and ln(p_{k+1/2} / p_{k-1/2}) and the mass flux all
// depend on the column's surface pressure, so the kernel takes the nodal
// ln ps as an input and evaluates them per column.  Pass A forms the column
// totals (mass-flux divergences, the geopotential sums, the ln ps tendency);
// pass B streams the levels with a (k, k+1) window and writes every nodal
// product the forward transform needs.  It replaces the XLA elementwise
// chains of
//   compute_diagnostic_state_hybrid        primitive_equations.py:1536-1602
//   _get_vertical_discretization_coeffs     :1483-1533
//   get_geopotential_diff_hybrid            :1629-1674
//   _get_pgf_lps_coefficient                :1826-1888
//   hybrid_vertical_advection               :1941-1957
//   PrimitiveEquationsHybrid._t_omega_over_p, curl_and_div_tendencies,
//     nodal_temperature_vertical_tendency, nodal_temperature_adiabatic_tendency,
//     nodal_log_pressure_tendency, explicit_terms          :2032-2362
//
// Input groups : Z D T TR[0..nt) U V GX GY LN      (LN = nodal ln ps)
// Output groups: CU CV KE TEMP UT VT LSP (TRN TU TV)[0..nt)
// KE carries the kinetic energy plus the nonlinear geopotential residual
// (geopotential of the full virtual temperature at the column's ps minus
// its linearisation at the reference state): both enter the divergence
// tendency as -laplacian(.), so one field carries their sum.  The output
// groups are those of the dry sigma kernel, so the sigma modal epilogue
// assembles the tendencies unchanged.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dino_common.h"

struct dino_nodal_hybrid {
  int K, I, J, nt, hum, ncloud, moist;
  int cloud[4];
  int iva;  // include vertical advection
  int tnc;  // reference temperature varies with level
  double R, kappa, gcr, hcr, ps_ref;
  double* c64;
  float* c32;
};

namespace dino {
namespace {

constexpr int HMAXTR = 6;

// constant block: sec2[J], cor[J], then per level: da, db, ah (K+1), bh (K+1),
// tref, alpha_ref, lr_ref, pgf_ref
struct HOff {
  int sec2, cor, da, db, ah, bh, tref, aref, lref, pref, total;
};

__host__ __device__ inline HOff hoff(int K, int J) {
  HOff o;
  o.sec2 = 0;
  o.cor = J;
  o.da = 2 * J;
  o.db = o.da + K;
  o.ah = o.db + K;
  o.bh = o.ah + K + 1;
  o.tref = o.bh + K + 1;
  o.aref = o.tref + K;
  o.lref = o.aref + K;
  o.pref = o.lref + K;
  o.total = o.pref + K;
  return o;
}

struct HGroups {
  int64_t z, d, t, tr[HMAXTR], u, v, gx, gy, ln;
  int64_t cu, cv, ke, te, ut, vt, lsp, trn[HMAXTR], tu[HMAXTR], tv[HMAXTR];
  int64_t n_in, n_out;
};

inline HGroups make_hgroups(int K, int nt, int64_t E) {
  HGroups g;
  std::memset(&g, 0, sizeof(g));
  int64_t o = 0;
  auto take = [&](int64_t lev) {
    int64_t r = o;
    o += lev * E;
    return r;
  };
  g.z = take(K);
  g.d = take(K);
  g.t = take(K);
  for (int i = 0; i < nt; ++i) g.tr[i] = take(K);
  g.u = take(K);
  g.v = take(K);
  g.gx = take(1);
  g.gy = take(1);
  g.ln = take(1);
  g.n_in = o;
  o = 0;
  g.cu = take(K);
  g.cv = take(K);
  g.ke = take(K);
  g.te = take(K);
  g.ut = take(K);
  g.vt = take(K);
  g.lsp = take(1);
  for (int i = 0; i < nt; ++i) {
    g.trn[i] = take(K);
    g.tu[i] = take(K);
    g.tv[i] = take(K);
  }
  g.n_out = o;
  return g;
}

struct HParams {
  int K, J, nt, hum, ncloud, iva, tnc, moist;
  int cloud[4];
  int64_t IJ;
  double R, kappa, gcr, hcr, inv_psref;
};

__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }

// Sum carried with its rounding error (Knuth's TwoSum per addition).  The
// geopotential needs suffix sums over the levels below k; the kernel walks
// down the column, so it forms them as column total minus prefix, and with
// both carried exactly to second order the difference keeps the digits a
// direct reverse sum would (the total is dominated by the reference
// temperature, the suffix near the surface is a few terms).
template <typename T>
struct TwoSumAcc {
  T hi = 0, lo = 0;
  __device__ __forceinline__ void add(T x) {
    // no multiply-add fusion into these additions: the error terms must be
    // those of the rounded summands, and the same in both passes
#pragma clang fp contract(off)
    const T s = hi + x;
    const T bp = s - hi;
    lo += (hi - (s - bp)) + (x - bp);
    hi = s;
  }
};

// a - b of two compensated sums
template <typename T>
__device__ __forceinline__ T twosum_diff(const TwoSumAcc<T>& a, const TwoSumAcc<T>& b) {
#pragma clang fp contract(off)
  return (a.hi - b.hi) + (a.lo - b.lo);
}

template <typename T>
__global__ __launch_bounds__(256) void nodal_hybrid_kernel(const T* __restrict__ in,
                                                           T* __restrict__ out,
                                                           const T* __restrict__ cst, HGroups g,
                                                           HParams p) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = blockIdx.y;
  if (col >= p.IJ) return;
  const int K = p.K;
  const HOff o = hoff(K, p.J);
  const int j = (int)(col % p.J);
  const int64_t IJ = p.IJ;
  auto X_ = [&](int64_t off, int lev, int k) -> T { return in[(off + e * lev + k) * IJ + col]; };
  auto O_ = [&](int64_t off, int lev, int k) -> T* { return out + (off + e * lev + k) * IJ + col; };
  const T s = cst[o.sec2 + j];
  const T fc = cst[o.cor + j];
  const T gx = X_(g.gx, 1, 0), gy = X_(g.gy, 1, 0);
  const T ps = dexp(X_(g.ln, 1, 0));
  const T floor_p = (T)1e-6 * ps;
  const T R = (T)p.R, kappa = (T)p.kappa;
  const bool moist = p.moist != 0, iva = p.iva != 0;
  // virtual-temperature factor of level k (primitive_equations.py:749-759)
  auto adjust = [&](int k) -> T {
    if (!moist) return (T)1;
    T a = (T)1 + (T)(p.gcr - 1) * X_(g.tr[p.hum], K, k);
    for (int c = 0; c < p.ncloud; ++c) a -= X_(g.tr[p.cloud[c]], K, k);
    return a;
  };
  // ---- pass A: column totals
  T TA = 0, TF = 0, LA = 0, LB = 0;
  TwoSumAcc<T> GS, GL;
  T p_up = cst[o.ah] + cst[o.bh] * ps;
  for (int k = 0; k < K; ++k) {
    const T u = X_(g.u, K, k), v = X_(g.v, K, k), d = X_(g.d, K, k), t = X_(g.t, K, k);
    const T da = cst[o.da + k], db = cst[o.db + k];
    const T udg = (u * gx + v * gy) * s;
    const T dp = da + db * ps;
    const T adv = db * ps * udg;
    TA += adv;
    TF += dp * d + adv;
    const T p_lo = cst[o.ah + k + 1] + cst[o.bh + k + 1] * ps;
    const T lr = dlog(fmax(p_lo, floor_p) / fmax(p_up, floor_p));
    GS.add((t + cst[o.tref + k]) * adjust(k) * lr);
    GL.add(cst[o.lref + k] * t);
    LA += d * da;
    LB += udg * db;
    p_up = p_lo;
  }
  *O_(g.lsp, 1, 0) = -(LA * ((T)1 / ps - (T)p.inv_psref) + LB);
  // ---- pass B
  T PA = 0, PF = 0;  // prefixes through level k
  TwoSumAcc<T> PGS, PGL;
  T aPA_m = 0, aPF_m = 0;              // alpha_ref_{k-1} * prefix_{k-1}
  T fu_m = 0, fv_m = 0, ft_m = 0, fx_m = 0, ftr_m[HMAXTR];
#pragma unroll
  for (int i = 0; i < HMAXTR; ++i) ftr_m[i] = 0;
  T u0 = X_(g.u, K, 0), v0 = X_(g.v, K, 0), t0 = X_(g.t, K, 0);
  T tr0[HMAXTR];
#pragma unroll
  for (int i = 0; i < HMAXTR; ++i) tr0[i] = i < p.nt ? X_(g.tr[i], K, 0) : (T)0;
  p_up = cst[o.ah] + cst[o.bh] * ps;
  for (int k = 0; k < K; ++k) {
    const bool last = (k == K - 1);
    T u1 = 0, v1 = 0, t1 = 0, tr1[HMAXTR];
#pragma unroll
    for (int i = 0; i < HMAXTR; ++i) tr1[i] = 0;
    if (!last) {
      u1 = X_(g.u, K, k + 1);
      v1 = X_(g.v, K, k + 1);
      t1 = X_(g.t, K, k + 1);
#pragma unroll
      for (int i = 0; i < HMAXTR; ++i)
        if (i < p.nt) tr1[i] = X_(g.tr[i], K, k + 1);
    }
    const T d = X_(g.d, K, k), z = X_(g.z, K, k);
    const T da = cst[o.da + k], db = cst[o.db + k];
    const T tref = cst[o.tref + k];
    const T udg = (u0 * gx + v0 * gy) * s;
    const T dp = da + db * ps;
    const T adv = db * ps * udg;
    PA += adv;
    PF += dp * d + adv;
    // Simmons-Burridge coefficients at this column's pressure
    const T p_lo = cst[o.ah + k + 1] + cst[o.bh + k + 1] * ps;
    const T lr = dlog(fmax(p_lo, floor_p) / fmax(p_up, floor_p));
    const T alpha = (T)1 - (p_up / dp) * lr;
    const T weight = lr * cst[o.bh + k] + alpha * db;
    const T dpf = fmax(dp, floor_p);
    // mass fluxes at the interface k + 1/2 (zero at the surface)
    const T bl = cst[o.bh + k + 1];
    const T Mf = last ? (T)0 : bl * TF - PF;
    const T Me = last ? (T)0 : bl * TA - PA;
    const T inv2dp = (T)-0.5 / dp;
    const T fu = Mf * (u1 - u0), fv = Mf * (v1 - v0), ft = Mf * (t1 - t0);
    const T vadv_u = (fu + fu_m) * inv2dp, vadv_v = (fv + fv_m) * inv2dp;
    const T vadv_t = (ft + ft_m) * inv2dp;
    T vadv_x = 0;
    if (p.tnc) {
      const T tref1 = last ? (T)0 : cst[o.tref + k + 1];
      const T fx = Me * (tref1 - tref);
      vadv_x = (fx + fx_m) * inv2dp;
      fx_m = fx;
    }
    const T adj = adjust(k);
    // pressure-gradient coefficient of grad ln ps, minus its reference part
    const T coef = R * (t0 + tref) / dpf * weight * ps * adj - cst[o.pref + k];
    // T omega / p with the reference alpha (primitive_equations.py:2032-2056)
    const T vgl = udg * ps * weight / dpf;
    const T aref = cst[o.aref + k];
    const T aPA = aref * PA, aPF = aref * PF;
    T tvar = t0;
    if (moist) {
      const T q = tr0[p.hum];
      const T den = (T)1 + (T)(p.hcr - 1) * q;
      tvar = t0 * (((T)1 + (T)(p.gcr - 1) * q) / den) + tref * (((T)(p.gcr - p.hcr) * q) / den);
    }
    const T mean_t = tref * (vgl - (aPA + aPA_m) / dp);
    const T var_t = tvar * (vgl - (aPF + aPF_m) / dp);
    const T temp = t0 * d + (iva ? vadv_t : (T)0) + vadv_x + kappa * (mean_t + var_t);
    const T zt = z + fc;
    *O_(g.cu, K, k) = -v0 * zt * s + ((iva ? -vadv_u : (T)0) + coef * gx) * s;
    *O_(g.cv, K, k) = u0 * zt * s + ((iva ? -vadv_v : (T)0) + coef * gy) * s;
    // geopotential of the full virtual temperature minus the reference
    // linearisation R [alpha_ref t + sum_{j > k} lr_ref_j t_j]
    const T tv_lr = (t0 + tref) * adj * lr;
    PGS.add(tv_lr);
    PGL.add(cst[o.lref + k] * t0);
    const T geo_full = R * (twosum_diff(GS, PGS) + alpha * (t0 + tref) * adj);
    const T geo_lin = R * (twosum_diff(GL, PGL) + aref * t0);
    *O_(g.ke, K, k) = (u0 * u0 + v0 * v0) * s / (T)2 + (geo_full - geo_lin);
    *O_(g.te, K, k) = temp;
    *O_(g.ut, K, k) = u0 * t0 * s;
    *O_(g.vt, K, k) = v0 * t0 * s;
#pragma unroll
    for (int i = 0; i < HMAXTR; ++i) {
      if (i < p.nt) {
        const T x0 = tr0[i];
        const T fx = Mf * (tr1[i] - x0);
        const T va = (fx + ftr_m[i]) * inv2dp;
        ftr_m[i] = fx;
        *O_(g.trn[i], K, k) = (iva ? va : (T)0) + x0 * d;
        *O_(g.tu[i], K, k) = u0 * x0 * s;
        *O_(g.tv[i], K, k) = v0 * x0 * s;
      }
    }
    fu_m = fu;
    fv_m = fv;
    ft_m = ft;
    aPA_m = aPA;
    aPF_m = aPF;
    p_up = p_lo;
    u0 = u1;
    v0 = v1;
    t0 = t1;
#pragma unroll
    for (int i = 0; i < HMAXTR; ++i) tr0[i] = tr1[i];
  }
}

template <typename T>
int launch_hybrid(const dino_nodal_hybrid* n, const void* in, void* out, int64_t E,
                  hipStream_t s) {
  HGroups g = make_hgroups(n->K, n->nt, E);
  HParams p;
  std::memset(&p, 0, sizeof(p));
  p.K = n->K;
  p.J = n->J;
  p.nt = n->nt;
  p.hum = n->hum;
  p.ncloud = n->ncloud;
  for (int i = 0; i < 4; ++i) p.cloud[i] = n->cloud[i];
  p.iva = n->iva;
  p.tnc = n->tnc;
  p.moist = n->moist;
  p.IJ = (int64_t)n->I * n->J;
  p.R = n->R;
  p.kappa = n->kappa;
  p.gcr = n->gcr;
  p.hcr = n->hcr;
  p.inv_psref = 1.0 / n->ps_ref;
  const T* cst;
  if constexpr (sizeof(T) == 8) cst = n->c64;
  else cst = n->c32;
  dim3 grid((unsigned)((p.IJ + 255) / 256), (unsigned)E);
  nodal_hybrid_kernel<T><<<grid, 256, 0, s>>>(static_cast<const T*>(in), static_cast<T*>(out),
                                              cst, g, p);
  DINO_LAUNCH_CHECK("hybrid nodal kernel");
  return DINO_OK;
}

}  // namespace
}  // namespace dino

using namespace dino;

extern "C" int dino_nodal_hybrid_create(int K, int I, int J, int n_tracers, int humidity_index,
                                        const int* cloud_indices, int n_clouds,
                                        int include_vertical_advection, const double* sec2,
                                        const double* coriolis, const double* a_boundaries,
                                        const double* b_boundaries, const double* tref,
                                        double ps_ref, double r_gas, double kappa,
                                        double r_vapor, double cp, double cp_vapor,
                                        dino_nodal_hybrid** out) {
  DINO_REQUIRE(out, DINO_EINVAL, "dino_nodal_hybrid_create: out is NULL");
  *out = nullptr;
  DINO_REQUIRE(K >= 1 && I >= 1 && J >= 1, DINO_EINVAL, "invalid nodal shape");
  DINO_REQUIRE(n_tracers >= 0 && n_tracers <= HMAXTR, DINO_EINVAL,
               "at most %d tracers are supported, got %d", HMAXTR, n_tracers);
  DINO_REQUIRE(n_clouds >= 0 && n_clouds <= 4, DINO_EINVAL, "at most 4 cloud tracers");
  DINO_REQUIRE(humidity_index < n_tracers, DINO_EINVAL, "humidity index out of range");
  DINO_REQUIRE(n_clouds == 0 || humidity_index >= 0, DINO_EINVAL,
               "cloud_keys requires humidity_key to be set.");
  DINO_REQUIRE(sec2 && coriolis && a_boundaries && b_boundaries && tref, DINO_EINVAL,
               "NULL constant array");
  DINO_REQUIRE(ps_ref > 0, DINO_EINVAL, "reference surface pressure must be positive");
  // reference-state coefficients: alpha, ln(p ratio) and the pgf coefficient
  // at ps_ref, with the same 1e-6 ps floor as the per-column ones
  std::vector<double> aref(K), lref(K), pref(K);
  for (int k = 0; k < K; ++k) {
    const double up = a_boundaries[k] + b_boundaries[k] * ps_ref;
    const double lo = a_boundaries[k + 1] + b_boundaries[k + 1] * ps_ref;
    const double dp = lo - up;
    DINO_REQUIRE(dp > 0, DINO_EINVAL, "Pressure must decrease monotonically with height.");
    const double fl = 1e-6 * ps_ref;
    const double lr = std::log(std::max(lo, fl) / std::max(up, fl));
    aref[k] = 1 - (up / dp) * lr;
    lref[k] = lr;
    const double db = b_boundaries[k + 1] - b_boundaries[k];
    const double w = lr * b_boundaries[k] + aref[k] * db;
    pref[k] = r_gas * tref[k] / std::max(dp, fl) * w * ps_ref;
  }
  dino_nodal_hybrid* n = new dino_nodal_hybrid();
  std::memset(n, 0, sizeof(*n));
  n->K = K;
  n->I = I;
  n->J = J;
  n->nt = n_tracers;
  n->hum = humidity_index;
  n->moist = humidity_index >= 0;
  n->ncloud = n_clouds;
  for (int i = 0; i < n_clouds; ++i) {
    if (cloud_indices[i] < 0 || cloud_indices[i] >= n_tracers) {
      delete n;
      set_error("cloud index out of range");
      return DINO_EINVAL;
    }
    n->cloud[i] = cloud_indices[i];
  }
  n->iva = include_vertical_advection;
  n->tnc = 0;
  for (int k = 1; k < K; ++k)
    if (tref[k] != tref[0]) n->tnc = 1;
  n->R = r_gas;
  n->kappa = kappa;
  n->gcr = r_vapor / r_gas;
  n->hcr = cp_vapor / cp;
  n->ps_ref = ps_ref;
  HOff o = hoff(K, J);
  std::vector<double> h(o.total, 0.0);
  for (int j = 0; j < J; ++j) {
    h[o.sec2 + j] = sec2[j];
    h[o.cor + j] = coriolis[j];
  }
  for (int k = 0; k <= K; ++k) {
    h[o.ah + k] = a_boundaries[k];
    h[o.bh + k] = b_boundaries[k];
  }
  for (int k = 0; k < K; ++k) {
    h[o.da + k] = a_boundaries[k + 1] - a_boundaries[k];
    h[o.db + k] = b_boundaries[k + 1] - b_boundaries[k];
    h[o.tref + k] = tref[k];
    h[o.aref + k] = aref[k];
    h[o.lref + k] = lref[k];
    h[o.pref + k] = pref[k];
  }
  std::vector<float> hf(h.begin(), h.end());
  hipError_t e = hipMalloc(&n->c64, sizeof(double) * o.total);
  if (e == hipSuccess) e = hipMalloc(&n->c32, sizeof(float) * o.total);
  if (e == hipSuccess)
    e = hipMemcpy(n->c64, h.data(), sizeof(double) * o.total, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(n->c32, hf.data(), sizeof(float) * o.total, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (n->c64) (void)hipFree(n->c64);
    if (n->c32) (void)hipFree(n->c32);
    delete n;
    return hip_fail(e, "dino_nodal_hybrid_create");
  }
  *out = n;
  return DINO_OK;
}

extern "C" int dino_nodal_hybrid_destroy(dino_nodal_hybrid* n) {
  if (n) {
    if (n->c64) (void)hipFree(n->c64);
    if (n->c32) (void)hipFree(n->c32);
    delete n;
  }
  return DINO_OK;
}

extern "C" int64_t dino_nodal_hybrid_fields(const dino_nodal_hybrid* n, int64_t members,
                                            int output) {
  if (!n || members < 0) return -1;
  HGroups g = make_hgroups(n->K, n->nt, members);
  return output ? g.n_out : g.n_in;
}

extern "C" int dino_nodal_hybrid_explicit(const dino_nodal_hybrid* n, const void* nodal_in,
                                          void* nodal_out, int64_t members, int dtype,
                                          void* stream) {
  DINO_REQUIRE(n && nodal_in && nodal_out, DINO_EINVAL,
               "dino_nodal_hybrid_explicit: NULL argument");
  DINO_REQUIRE(members >= 1 && members <= 65535, DINO_EINVAL, "members must be in [1, 65535]");
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "unsupported dtype");
  if (dtype == DINO_F64)
    return launch_hybrid<double>(n, nodal_in, nodal_out, members, as_stream(stream));
  return launch_hybrid<float>(n, nodal_in, nodal_out, members, as_stream(stream));
}
