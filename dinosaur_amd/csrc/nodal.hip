// Fused nodal nonlinear terms of PrimitiveEquationsSigma.explicit_terms and
// their adjoint.
//
// One thread owns one (member, longitude, latitude) column and walks the K
// levels twice: pass A forms the column totals of u.grad(ln ps) dsigma and
// (div + u.grad(ln ps)) dsigma (needed by sigma-dot at every interface) and
// the moist geopotential total; pass B streams the levels with a three-level
// window, producing sigma-dot, centred vertical advection, the T omega / p
// terms and every nodal product the forward transform needs.  All loads and
// stores are coalesced along latitude.  It replaces the chain of XLA
// elementwise ops and cumsums in
//   compute_diagnostic_state_sigma          primitive_equations.py:168-225
//   curl_and_div_tendencies                 :1001-1059
//   kinetic_energy_tendency                 :778-785
//   horizontal_scalar_advection (nodal)     :795-805
//   nodal_temperature_vertical_tendency     :1061-1079
//   nodal_temperature_adiabatic_tendency    :1081-1115 (_t_omega_over_sigma_sp :965-999)
//   nodal_log_pressure_tendency             :1117-1124
//   humidity terms                          :861-963
//   centered_vertical_advection             sigma_coordinates.py:344-405
//
// Field groups live in one (fields, I, J) buffer each; group g of a member e
// and level k is field off[g] + e * levels(g) + k.
//
// Input groups : Z D T TR[0..nt) U V GX GY [LP QX QY]
// Output groups: CU CV KE TEMP UT VT LSP (TRN TU TV)[0..nt) [CURL DC]
// For moist equations KE holds the kinetic energy plus the virtual-temperature
// geopotential: both enter the divergence tendency as -laplacian(.)
// (primitive_equations.py:1001-1059, 1126-1194), so one field carries their
// sum through the forward transform instead of two.
#include <cstring>

#include "dino_common.h"

struct dino_nodal {
  int K, I, J, nt, hum, ncloud;
  int cloud[4];
  int iva;       // include vertical advection
  int tnc;       // reference temperature varies with level
  int moist;
  double R, kappa, gcr, hcr, rv_minus_r, gcr1;
  // device constants (FP64 and FP32 copies): sec2[J], cor[J], dsig[K],
  // c2c[K] (1 / center_to_center, last unused), ss[K] (cumsum dsigma),
  // alpha[K], tref[K], gp[K] (geopotential off-diagonal weights, gp[0] = 0),
  // gd[K] (geopotential diagonal R alpha)
  double* c64;
  float* c32;
};

namespace dino {
namespace {

enum { C_SEC2 = 0, C_COR = 1 };  // followed by per-level arrays
struct COff {
  int sec2, cor, dsig, c2c, ss, alpha, tref, gp, gd, total;
};

__host__ __device__ inline COff coff(int K, int J) {
  COff o;
  o.sec2 = 0;
  o.cor = J;
  o.dsig = 2 * J;
  o.c2c = o.dsig + K;
  o.ss = o.c2c + K;
  o.alpha = o.ss + K;
  o.tref = o.alpha + K;
  o.gp = o.tref + K;
  o.gd = o.gp + K;
  o.total = o.gd + K;
  return o;
}

constexpr int MAXTR = 6;

struct Groups {
  // input group offsets (in fields of I*J) for member stride computation
  int64_t z, d, t, tr[MAXTR], u, v, gx, gy, lp, qx, qy;
  // output group offsets
  int64_t cu, cv, ke, te, ut, vt, lsp, trn[MAXTR], tu[MAXTR], tv[MAXTR], curl, dc;
  int64_t n_in, n_out;
};

inline Groups make_groups(int K, int nt, int moist, int64_t E) {
  Groups g;
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
  if (moist) {
    g.lp = take(1);
    g.qx = take(K);
    g.qy = take(K);
  }
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
  if (moist) {
    g.curl = take(K);
    g.dc = take(K);
  }
  g.n_out = o;
  return g;
}

struct Params {
  int K, J, nt, hum, ncloud, iva, tnc, moist;
  int cloud[4];
  int64_t IJ;
  int64_t E;
  double R, kappa, gcr, hcr, rvr, gcr1;
};

// ---------------------------------------------------------------- forward
template <typename T>
__global__ __launch_bounds__(256) void nodal_forward_kernel(
    const T* __restrict__ in, T* __restrict__ out, const T* __restrict__ cst, Groups g,
    Params p) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = blockIdx.y;
  if (col >= p.IJ) return;
  const int K = p.K;
  const COff o = coff(K, p.J);
  const int j = (int)(col % p.J);
  const int64_t IJ = p.IJ;
  // field (group offset, levels) of member e at level k
  auto I_ = [&](int64_t off, int lev, int k) -> const T* {
    return in + (off + e * lev + k) * IJ + col;
  };
  auto O_ = [&](int64_t off, int lev, int k) -> T* { return out + (off + e * lev + k) * IJ + col; };
  const T s = cst[o.sec2 + j];
  const T fc = cst[o.cor + j];
  const T gx = *I_(g.gx, 1, 0);
  const T gy = *I_(g.gy, 1, 0);
  const T R = (T)p.R, kappa = (T)p.kappa;
  const bool moist = p.moist != 0;
  const T lp = moist ? *I_(g.lp, 1, 0) : (T)0;
  // ---- pass A: column totals
  T TE = 0, TF = 0, GT = 0;
  for (int k = 0; k < K; ++k) {
    T u = *I_(g.u, K, k), v = *I_(g.v, K, k), d = *I_(g.d, K, k);
    T A = (u * gx + v * gy) * s;
    T ds = cst[o.dsig + k];
    TE += A * ds;
    TF += (d + A) * ds;
    if (moist && k >= 1) {
      T q = *I_(g.tr[p.hum], K, k), t = *I_(g.t, K, k);
      T td = q * (t + cst[o.tref + k]) * (T)p.gcr1;
      GT += cst[o.gp + k] * td;
    }
  }
  *O_(g.lsp, 1, 0) = -TE;
  // ---- pass B: stream levels with windows (k-1, k, k+1)
  T PE = 0, PF = 0, PEm = 0, PFm = 0;  // prefixes at k and k-1
  T Wm = 0, Xm = 0;                     // sigma-dot at interface k-1/2
  // flux of the previous interface for each advected field: U, V, T, tracers
  T fu_m = 0, fv_m = 0, ft_m = 0, fx_m = 0, ftr_m[MAXTR];
#pragma unroll
  for (int i = 0; i < MAXTR; ++i) ftr_m[i] = 0;
  T u0 = *I_(g.u, K, 0), v0 = *I_(g.v, K, 0), t0 = *I_(g.t, K, 0);
  T tr0[MAXTR];
#pragma unroll
  for (int i = 0; i < MAXTR; ++i) tr0[i] = i < p.nt ? *I_(g.tr[i], K, 0) : (T)0;
  T GP = 0;  // prefix sum_{1 <= h <= k} gp[h] td[h]
  for (int k = 0; k < K; ++k) {
    const bool last = (k == K - 1);
    T u1 = 0, v1 = 0, t1 = 0, tr1[MAXTR];
#pragma unroll
    for (int i = 0; i < MAXTR; ++i) tr1[i] = 0;
    if (!last) {
      u1 = *I_(g.u, K, k + 1);
      v1 = *I_(g.v, K, k + 1);
      t1 = *I_(g.t, K, k + 1);
#pragma unroll
      for (int i = 0; i < MAXTR; ++i)
        if (i < p.nt) tr1[i] = *I_(g.tr[i], K, k + 1);
    }
    const T d = *I_(g.d, K, k), z = *I_(g.z, K, k);
    const T ds = cst[o.dsig + k];
    const T A = (u0 * gx + v0 * gy) * s;
    PE += A * ds;
    PF += (d + A) * ds;
    T W = 0, X = 0;
    if (!last) {
      W = cst[o.ss + k] * TF - PF;
      X = cst[o.ss + k] * TE - PE;
    }
    const T c = last ? (T)0 : cst[o.c2c + k];
    // vertical advection -(w dx/dsigma), averaged over the two interfaces
    T fu = W * (u1 - u0) * c, fv = W * (v1 - v0) * c, ft = W * (t1 - t0) * c;
    T vadv_u = (T)-0.5 * (fu + fu_m), vadv_v = (T)-0.5 * (fv + fv_m);
    T vadv_t = (T)-0.5 * (ft + ft_m);
    T tref = cst[o.tref + k];
    T vadv_x = 0;
    if (p.tnc) {
      T tref1 = last ? (T)0 : cst[o.tref + k + 1];
      T fx = last ? (T)0 : X * (tref1 - tref) * c;
      vadv_x = (T)-0.5 * (fx + fx_m);
      fx_m = fx;
    }
    // humidity
    T q = 0, adj = 1;
    if (moist) {
      q = tr0[p.hum];
      adj = (T)1 + (T)(p.gcr - 1) * q;
      for (int c2 = 0; c2 < p.ncloud; ++c2) adj -= tr0[p.cloud[c2]];
    }
    // T omega / p terms
    const T al = cst[o.alpha + k];
    const T alm = k > 0 ? cst[o.alpha + k - 1] : (T)0;
    const T gpE = (al * PE + alm * PEm) / ds;
    const T gpF = (al * PF + alm * PFm) / ds;
    T tv;
    if (!moist) {
      tv = t0;
    } else {
      T den = (T)1 + (T)(p.hcr - 1) * q;
      tv = t0 * (((T)1 + (T)(p.gcr - 1) * q) / den) + tref * (((T)(p.gcr - p.hcr) * q) / den);
    }
    const T mean_t = tref * (A - gpE);
    const T var_t = tv * (A - gpF);
    const bool iva = p.iva != 0;
    T temp = t0 * d + (iva ? vadv_t : (T)0) + vadv_x + kappa * (mean_t + var_t);
    const T rt = R * t0 * adj;
    const T zt = z + fc;
    *O_(g.cu, K, k) = -v0 * zt * s + ((iva ? -vadv_u : (T)0) + rt * gx) * s;
    *O_(g.cv, K, k) = u0 * zt * s + ((iva ? -vadv_v : (T)0) + rt * gy) * s;
    const T kinetic = (u0 * u0 + v0 * v0) * s / (T)2;
    if (!moist) *O_(g.ke, K, k) = kinetic;
    *O_(g.te, K, k) = temp;
    *O_(g.ut, K, k) = u0 * t0 * s;
    *O_(g.vt, K, k) = v0 * t0 * s;
#pragma unroll
    for (int i = 0; i < MAXTR; ++i) {
      if (i < p.nt) {
        T x0 = tr0[i];
        T fx = W * (tr1[i] - x0) * c;
        T va = (T)-0.5 * (fx + ftr_m[i]);
        ftr_m[i] = fx;
        *O_(g.trn[i], K, k) = (iva ? va : (T)0) + x0 * d;
        *O_(g.tu[i], K, k) = u0 * x0 * s;
        *O_(g.tv[i], K, k) = v0 * x0 * s;
      }
    }
    if (moist) {
      T qx = *I_(g.qx, K, k), qy = *I_(g.qy, K, k);
      T cf = tref * (T)p.rvr;
      *O_(g.curl, K, k) = cf * s * (gx * qy - gy * qx);
      *O_(g.dc, K, k) = cf * s * (qx * gx + qy * gy) + q * lp * tref * (T)p.rvr;
      T td = q * (t0 + tref) * (T)p.gcr1;
      if (k >= 1) GP += cst[o.gp + k] * td;
      *O_(g.ke, K, k) = kinetic + (cst[o.gd + k] * td + (GT - GP));
    }
    fu_m = fu;
    fv_m = fv;
    ft_m = ft;
    PEm = PE;
    PFm = PF;
    Wm = W;
    Xm = X;
    u0 = u1;
    v0 = v1;
    t0 = t1;
#pragma unroll
    for (int i = 0; i < MAXTR; ++i) tr0[i] = tr1[i];
  }
  (void)Wm;
  (void)Xm;
}

// ---------------------------------------------------------------- adjoint
// Reads the forward inputs and the output cotangents, writes the input
// cotangents.  Pass 1 (down the column) recomputes the totals TE, TF and the
// cotangents of those totals (sum ss_k Wbar_k, sum ss_k Xbar_k, -lsp bar) and
// the moist geopotential cotangent total; pass 2 (up the column) carries the
// suffix sums that are the adjoints of the downward prefix sums.
// Levels are streamed through a register window: pass 1 keeps levels
// (k, k + 1) for the interface terms, pass 2 (upwards) keeps k + 1, k and a
// prefetched k - 1, so every value is loaded once per pass.  NTR is the
// tracer count (the humidity and cloud indices select among them with
// unrolled compares, never a runtime register index).
template <typename T, int NTR>
__global__ __launch_bounds__(256) void nodal_adjoint_kernel(
    const T* __restrict__ in, const T* __restrict__ gout, T* __restrict__ gin,
    const T* __restrict__ cst, Groups g, Params p) {
  constexpr int NA = NTR > 0 ? NTR : 1;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = blockIdx.y;
  if (col >= p.IJ) return;
  const int K = p.K;
  const COff o = coff(K, p.J);
  const int j = (int)(col % p.J);
  const int64_t IJ = p.IJ;
  auto X_ = [&](int64_t off, int lev, int k) -> T {
    return in[(off + e * lev + k) * IJ + col];
  };
  auto B_ = [&](int64_t off, int lev, int k) -> T {
    return gout[(off + e * lev + k) * IJ + col];
  };
  auto G_ = [&](int64_t off, int lev, int k) -> T& {
    return gin[(off + e * lev + k) * IJ + col];
  };
  const T s = cst[o.sec2 + j];
  const T fc = cst[o.cor + j];
  const T gx = X_(g.gx, 1, 0), gy = X_(g.gy, 1, 0);
  const T R = (T)p.R, kappa = (T)p.kappa;
  const bool moist = p.moist != 0, iva = p.iva != 0;
  const T lp = moist ? X_(g.lp, 1, 0) : (T)0;
  const int hum = p.hum;

  // per-level values: "light" = what the interface terms of a neighbour need
  struct XL {
    T u, v, t, tr[NA];
  };
  struct BL {
    T cu, cv, te, trn[NA];
  };
  struct XF {
    XL l;
    T d, z, qx, qy;
  };
  struct BF {
    BL l;
    T ke, ut, vt, tu[NA], tv[NA], curl, dc, geo;
  };
  auto load_xl = [&](int k, XL& x) {
    x.u = X_(g.u, K, k);
    x.v = X_(g.v, K, k);
    x.t = X_(g.t, K, k);
#pragma unroll
    for (int i = 0; i < NTR; ++i) x.tr[i] = X_(g.tr[i], K, k);
  };
  auto load_bl = [&](int k, BL& b) {
    b.cu = B_(g.cu, K, k);
    b.cv = B_(g.cv, K, k);
    b.te = B_(g.te, K, k);
#pragma unroll
    for (int i = 0; i < NTR; ++i) b.trn[i] = B_(g.trn[i], K, k);
  };
  auto load_xf = [&](int k, XF& x) {
    load_xl(k, x.l);
    x.d = X_(g.d, K, k);
    x.z = X_(g.z, K, k);
    x.qx = moist ? X_(g.qx, K, k) : (T)0;
    x.qy = moist ? X_(g.qy, K, k) : (T)0;
  };
  auto load_bf = [&](int k, BF& b) {
    load_bl(k, b.l);
    b.ke = B_(g.ke, K, k);
    b.ut = B_(g.ut, K, k);
    b.vt = B_(g.vt, K, k);
#pragma unroll
    for (int i = 0; i < NTR; ++i) {
      b.tu[i] = B_(g.tu[i], K, k);
      b.tv[i] = B_(g.tv[i], K, k);
    }
    b.curl = moist ? B_(g.curl, K, k) : (T)0;
    b.dc = moist ? B_(g.dc, K, k) : (T)0;
    b.geo = moist ? b.ke : (T)0;  // KE carries the geopotential too
  };
  // cotangents of vadv(W, x) at a level
  auto vb_u = [&](const BL& b) -> T { return iva ? -s * b.cu : (T)0; };
  auto vb_v = [&](const BL& b) -> T { return iva ? -s * b.cv : (T)0; };
  auto vb_t = [&](const BL& b) -> T { return iva ? b.te : (T)0; };
  auto vb_tr = [&](const BL& b, int i) -> T { return iva ? b.trn[i] : (T)0; };
  // Wbar at interface k = sum_x Fbar_x (x_{k+1} - x_k) c_k  (levels k, k + 1)
  auto wbar = [&](int k, const XL& x0, const XL& x1, const BL& b0, const BL& b1) -> T {
    T c = cst[o.c2c + k];
    T acc = 0;
    acc += (T)-0.5 * (vb_u(b0) + vb_u(b1)) * (x1.u - x0.u);
    acc += (T)-0.5 * (vb_v(b0) + vb_v(b1)) * (x1.v - x0.v);
    acc += (T)-0.5 * (vb_t(b0) + vb_t(b1)) * (x1.t - x0.t);
#pragma unroll
    for (int i = 0; i < NTR; ++i)
      acc += (T)-0.5 * (vb_tr(b0, i) + vb_tr(b1, i)) * (x1.tr[i] - x0.tr[i]);
    return acc * c;
  };
  auto xbar = [&](int k, const BL& b0, const BL& b1) -> T {
    if (!p.tnc) return (T)0;
    T c = cst[o.c2c + k];
    return (T)-0.5 * (b0.te + b1.te) * (cst[o.tref + k + 1] - cst[o.tref + k]) * c;
  };
  // humidity and the cloud sum of a level (unrolled selects)
  auto hum_of = [&](const XL& x) -> T {
    T q = 0;
#pragma unroll
    for (int i = 0; i < NTR; ++i) q = (i == hum) ? x.tr[i] : q;
    return q;
  };
  auto is_cloud = [&](int i) -> bool {
    bool r = false;
    for (int c2 = 0; c2 < p.ncloud; ++c2) r = r || (p.cloud[c2] == i);
    return r;
  };
  auto tv_of = [&](const XL& x, int k) -> T {
    if (!moist) return x.t;
    T qq = hum_of(x);
    T dn = (T)1 + (T)(p.hcr - 1) * qq;
    return x.t * (((T)1 + (T)(p.gcr - 1) * qq) / dn) +
           cst[o.tref + k] * (((T)(p.gcr - p.hcr) * qq) / dn);
  };

  // ---- pass 1: totals and the cotangents of the totals
  T TE = 0, TF = 0, TFb = 0, TEb = -B_(g.lsp, 1, 0), GEOT = 0;
  {
    XL xa, xb;
    BL ba, bb;
    load_xl(0, xa);
    load_bl(0, ba);
    for (int k = 0; k < K; ++k) {
      const int kn = k + 1 < K ? k + 1 : k;
      load_xl(kn, xb);
      load_bl(kn, bb);
      const T d = X_(g.d, K, k);
      const T A = (xa.u * gx + xa.v * gy) * s;
      const T ds = cst[o.dsig + k];
      TE += A * ds;
      TF += (d + A) * ds;
      if (k < K - 1) {
        TFb += cst[o.ss + k] * wbar(k, xa, xb, ba, bb);
        TEb += cst[o.ss + k] * xbar(k, ba, bb);
      }
      if (moist) GEOT += B_(g.ke, K, k);
      xa = xb;
      ba = bb;
    }
  }
  // ---- pass 2: k = K-1 .. 0
  T gxb = 0, gyb = 0, lpb = 0;
  T SPF = 0, SPE = 0;  // suffix sums of PFbar, PEbar over k' >= k
  T SB = 0, SE = 0;    // suffix sums of b, e over k' > k (forward values)
  T SGEO = 0;          // suffix sum of geo bars over k' >= k
  XF xc, xn;
  BF bc, bn;
  XL xu = {};
  BL bu = {};
  T tv_up = 0;
  load_xf(K - 1, xc);
  load_bf(K - 1, bc);
  for (int k = K - 1; k >= 0; --k) {
    const int kd = k > 0 ? k - 1 : k;  // level k - 1, prefetched for the next step
    load_xf(kd, xn);
    load_bf(kd, bn);
    const T ds = cst[o.dsig + k];
    const T u = xc.l.u, v = xc.l.v, d = xc.d, t = xc.l.t;
    const T z = xc.z;
    const T A = (u * gx + v * gy) * s;
    const T b = (d + A) * ds, ev = A * ds;
    const T PF = TF - SB, PE = TE - SE;    // prefixes at k
    const T PFm = PF - b, PEm = PE - ev;  // prefixes at k-1
    const T tref = cst[o.tref + k];
    const T al = cst[o.alpha + k];
    const T alm = k > 0 ? cst[o.alpha + k - 1] : (T)0;
    // forward sigma-dot at interfaces k and k-1
    const T W = (k < K - 1) ? cst[o.ss + k] * TF - PF : (T)0;
    const T Wm = (k > 0) ? cst[o.ss + k - 1] * TF - PFm : (T)0;
    // humidity quantities at k
    T q = 0, adj = 1, den = 1;
    if (moist) {
      q = hum_of(xc.l);
      adj = (T)1 + (T)(p.gcr - 1) * q;
#pragma unroll
      for (int i = 0; i < NTR; ++i)
        if (is_cloud(i)) adj -= xc.l.tr[i];
      den = (T)1 + (T)(p.hcr - 1) * q;
    }
    const T tv = tv_of(xc.l, k);
    // output cotangents at k
    const T cu = bc.l.cu, cv = bc.l.cv, ke = bc.ke;
    const T te = bc.l.te, ut = bc.ut, vt = bc.vt;
    const T om = kappa * te;  // cotangent of both T omega / p terms
    // prefix cotangents from T omega / p at levels k and k+1
    T pfb = -al * tv * om / ds;
    T peb = -al * tref * om / ds;
    if (k < K - 1) {
      T ds1 = cst[o.dsig + k + 1];
      T om1 = kappa * bu.te;
      pfb += -al * tv_up * om1 / ds1;
      peb += -al * cst[o.tref + k + 1] * om1 / ds1;
      pfb += -wbar(k, xc.l, xu, bc.l, bu);
      peb += -xbar(k, bc.l, bu);
    }
    SPF += pfb;
    SPE += peb;
    const T bbar = SPF + TFb;
    const T ebar = SPE + TEb;
    const T gpF = (al * PF + alm * PFm) / ds;
    // ---- input cotangents at level k
    T zb = 0, db = 0, tb = 0, ub = 0, vb = 0, qb = 0;
    const T zt = z + fc;
    // CU, CV
    zb += -v * s * cu + u * s * cv;
    vb += -zt * s * cu;
    ub += zt * s * cv;
    const T rt_bar_cu = R * adj * gx * s * cu + R * adj * gy * s * cv;  // d/dT
    tb += rt_bar_cu;
    const T adjb = R * t * (gx * s * cu + gy * s * cv);
    gxb += R * t * adj * s * cu;
    gyb += R * t * adj * s * cv;
    // KE
    ub += u * s * ke;
    vb += v * s * ke;
    // TEMP direct
    tb += d * te;
    db += t * te;
    // UT, VT
    ub += t * s * ut;
    tb += u * s * ut;
    vb += t * s * vt;
    tb += v * s * vt;
    // T omega / p: Abar and Tv bar
    T Abar = ds * (bbar + ebar) + tref * om + tv * om;
    const T tvb = (A - gpF) * om;
    if (!moist) {
      tb += tvb;
    } else {
      tb += (((T)1 + (T)(p.gcr - 1) * q) / den) * tvb;
      qb += (t + tref) * ((T)(p.gcr - p.hcr) / (den * den)) * tvb;
    }
    db += ds * bbar;
    ub += gx * s * Abar;
    vb += gy * s * Abar;
    gxb += u * s * Abar;
    gyb += v * s * Abar;
    // vertical advection cotangents on the advected fields
    // x_k gets Fbar_k * (-W c_k) (interface k) and Fbar_{k-1} * (W_{k-1} c_{k-1})
    T c = (k < K - 1) ? cst[o.c2c + k] : (T)0;
    T cm = (k > 0) ? cst[o.c2c + k - 1] : (T)0;
    auto vadv_bar = [&](T vbk, T vbk1, T vbkm) -> T {
      T r = 0;
      if (k < K - 1) r += (T)-0.5 * (vbk + vbk1) * (-W * c);
      if (k > 0) r += (T)-0.5 * (vbkm + vbk) * (Wm * cm);
      return r;
    };
    const bool hu = k < K - 1, hd = k > 0;
    ub += vadv_bar(vb_u(bc.l), hu ? vb_u(bu) : (T)0, hd ? vb_u(bn.l) : (T)0);
    vb += vadv_bar(vb_v(bc.l), hu ? vb_v(bu) : (T)0, hd ? vb_v(bn.l) : (T)0);
    tb += vadv_bar(vb_t(bc.l), hu ? vb_t(bu) : (T)0, hd ? vb_t(bn.l) : (T)0);
    // tracers
    T trb[NA];
#pragma unroll
    for (int i = 0; i < NTR; ++i) {
      T x = xc.l.tr[i];
      T trn = bc.l.trn[i], tu = bc.tu[i], tvv = bc.tv[i];
      T xb = d * trn + u * s * tu + v * s * tvv;
      db += x * trn;
      ub += x * s * tu;
      vb += x * s * tvv;
      xb += vadv_bar(vb_tr(bc.l, i), hu ? vb_tr(bu, i) : (T)0, hd ? vb_tr(bn.l, i) : (T)0);
      trb[i] = xb;
    }
    if (moist) {
      // virtual-temperature adjustment of R T
      qb += (T)(p.gcr - 1) * adjb;
#pragma unroll
      for (int i = 0; i < NTR; ++i)
        if (is_cloud(i)) trb[i] -= adjb;
      T qx = xc.qx, qy = xc.qy;
      T curl = bc.curl, dcb = bc.dc, geob = bc.geo;
      T cf = tref * (T)p.rvr;
      gxb += cf * s * (qy * curl + qx * dcb);
      gyb += cf * s * (-qx * curl + qy * dcb);
      G_(g.qx, K, k) = cf * s * (-gy * curl + gx * dcb);
      G_(g.qy, K, k) = cf * s * (gx * curl + gy * dcb);
      qb += lp * tref * (T)p.rvr * dcb;
      lpb += q * tref * (T)p.rvr * dcb;
      // geopotential: tdbar_k = gd_k geo_k + gp_k sum_{k' < k} geo_k'
      SGEO += geob;
      T tdb = cst[o.gd + k] * geob + (k >= 1 ? cst[o.gp + k] * (GEOT - SGEO) : (T)0);
      qb += (t + tref) * (T)p.gcr1 * tdb;
      tb += q * (T)p.gcr1 * tdb;
#pragma unroll
      for (int i = 0; i < NTR; ++i)
        if (i == hum) trb[i] += qb;
    }
#pragma unroll
    for (int i = 0; i < NTR; ++i) G_(g.tr[i], K, k) = trb[i];
    G_(g.z, K, k) = zb;
    G_(g.d, K, k) = db;
    G_(g.t, K, k) = tb;
    G_(g.u, K, k) = ub;
    G_(g.v, K, k) = vb;
    SB += b;
    SE += ev;
    // slide the window down one level
    xu = xc.l;
    bu = bc.l;
    tv_up = tv;
    xc = xn;
    bc = bn;
  }
  G_(g.gx, 1, 0) = gxb;
  G_(g.gy, 1, 0) = gyb;
  if (moist) G_(g.lp, 1, 0) = lpb;
}

template <typename T>
int launch(bool adjoint, const dino_nodal* n, const void* in, const void* gout, void* out,
           int64_t E, hipStream_t s) {
  Groups g = make_groups(n->K, n->nt, n->moist, E);
  Params p;
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
  p.E = E;
  p.R = n->R;
  p.kappa = n->kappa;
  p.gcr = n->gcr;
  p.hcr = n->hcr;
  p.rvr = n->rv_minus_r;
  p.gcr1 = n->gcr1;
  const T* cst;
  if constexpr (sizeof(T) == 8) cst = n->c64;
  else cst = n->c32;
  dim3 grid((unsigned)((p.IJ + 255) / 256), (unsigned)E);
  if (!adjoint)
    nodal_forward_kernel<T><<<grid, 256, 0, s>>>(static_cast<const T*>(in), static_cast<T*>(out),
                                                 cst, g, p);
  else
    switch (n->nt) {
#define ADJ_CASE(NTR)                                                                 \
  case NTR:                                                                           \
    nodal_adjoint_kernel<T, NTR><<<grid, 256, 0, s>>>(static_cast<const T*>(in),      \
                                                      static_cast<const T*>(gout),    \
                                                      static_cast<T*>(out), cst, g, p); \
    break;
      ADJ_CASE(0) ADJ_CASE(1) ADJ_CASE(2) ADJ_CASE(3) ADJ_CASE(4) ADJ_CASE(5) ADJ_CASE(6)
#undef ADJ_CASE
      default:
        set_error("nodal kernel supports up to %d tracers", MAXTR);
        return DINO_EINVAL;
    }
  DINO_LAUNCH_CHECK("nodal kernel");
  return DINO_OK;
}

}  // namespace
}  // namespace dino

using namespace dino;

extern "C" int dino_nodal_create(int K, int I, int J, int n_tracers, int humidity_index,
                                 const int* cloud_indices, int n_clouds,
                                 int include_vertical_advection, const double* sec2,
                                 const double* coriolis, const double* dsigma,
                                 const double* c2c_inv, const double* sum_sigma,
                                 const double* alpha, const double* tref, const double* geo_off,
                                 const double* geo_diag, double r_gas, double kappa,
                                 double r_vapor, double cp, double cp_vapor,
                                 dino_nodal** out) {
  DINO_REQUIRE(out, DINO_EINVAL, "dino_nodal_create: out is NULL");
  *out = nullptr;
  DINO_REQUIRE(K >= 1 && I >= 1 && J >= 1, DINO_EINVAL, "invalid nodal shape");
  DINO_REQUIRE(n_tracers >= 0 && n_tracers <= MAXTR, DINO_EINVAL,
               "at most %d tracers are supported, got %d", MAXTR, n_tracers);
  DINO_REQUIRE(n_clouds >= 0 && n_clouds <= 4, DINO_EINVAL, "at most 4 cloud tracers");
  DINO_REQUIRE(humidity_index < n_tracers, DINO_EINVAL, "humidity index out of range");
  DINO_REQUIRE(n_clouds == 0 || humidity_index >= 0, DINO_EINVAL,
               "cloud_keys requires humidity_key to be set.");
  DINO_REQUIRE(sec2 && coriolis && dsigma && c2c_inv && sum_sigma && alpha && tref, DINO_EINVAL,
               "NULL constant array");
  dino_nodal* n = new dino_nodal();
  std::memset(n, 0, sizeof(*n));
  n->K = K;
  n->I = I;
  n->J = J;
  n->nt = n_tracers;
  n->hum = humidity_index;
  n->moist = humidity_index >= 0;
  n->ncloud = n_clouds;
  for (int i = 0; i < n_clouds; ++i) {
    DINO_REQUIRE(cloud_indices[i] >= 0 && cloud_indices[i] < n_tracers, DINO_EINVAL,
                 "cloud index out of range");
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
  n->rv_minus_r = r_vapor - r_gas;
  n->gcr1 = r_vapor / r_gas - 1;
  COff o = coff(K, J);
  std::vector<double> h(o.total, 0.0);
  for (int j = 0; j < J; ++j) {
    h[o.sec2 + j] = sec2[j];
    h[o.cor + j] = coriolis[j];
  }
  for (int k = 0; k < K; ++k) {
    h[o.dsig + k] = dsigma[k];
    h[o.c2c + k] = k < K - 1 ? c2c_inv[k] : 0.0;
    h[o.ss + k] = k < K - 1 ? sum_sigma[k] : 0.0;
    h[o.alpha + k] = alpha[k];
    h[o.tref + k] = tref[k];
    h[o.gp + k] = geo_off ? geo_off[k] : 0.0;
    h[o.gd + k] = geo_diag ? geo_diag[k] : 0.0;
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
    return hip_fail(e, "dino_nodal_create");
  }
  *out = n;
  return DINO_OK;
}

extern "C" int dino_nodal_destroy(dino_nodal* n) {
  if (n) {
    if (n->c64) (void)hipFree(n->c64);
    if (n->c32) (void)hipFree(n->c32);
    delete n;
  }
  return DINO_OK;
}

extern "C" int64_t dino_nodal_fields(const dino_nodal* n, int64_t members, int output) {
  if (!n || members < 0) return -1;
  Groups g = make_groups(n->K, n->nt, n->moist, members);
  return output ? g.n_out : g.n_in;
}

extern "C" int dino_nodal_explicit(const dino_nodal* n, const void* nodal_in, void* nodal_out,
                                   int64_t members, int dtype, void* stream) {
  DINO_REQUIRE(n && nodal_in && nodal_out, DINO_EINVAL, "dino_nodal_explicit: NULL argument");
  DINO_REQUIRE(members >= 1 && members <= 65535, DINO_EINVAL, "members must be in [1, 65535]");
  if (dtype == DINO_F64)
    return launch<double>(false, n, nodal_in, nullptr, nodal_out, members, as_stream(stream));
  return launch<float>(false, n, nodal_in, nullptr, nodal_out, members, as_stream(stream));
}

extern "C" int dino_nodal_explicit_adjoint(const dino_nodal* n, const void* nodal_in,
                                           const void* out_cot, void* in_cot, int64_t members,
                                           int dtype, void* stream) {
  DINO_REQUIRE(n && nodal_in && out_cot && in_cot, DINO_EINVAL,
               "dino_nodal_explicit_adjoint: NULL argument");
  DINO_REQUIRE(members >= 1 && members <= 65535, DINO_EINVAL, "members must be in [1, 65535]");
  if (dtype == DINO_F64)
    return launch<double>(true, n, nodal_in, out_cot, in_cot, members, as_stream(stream));
  return launch<float>(true, n, nodal_in, out_cot, in_cot, members, as_stream(stream));
}
