// Fused modal-space stages of PrimitiveEquationsSigma.explicit_terms and
// their adjoints.
//
// prep:     state (vorticity, divergence, temperature, tracers, ln ps) ->
//           the one batched inverse-transform input: copies of the prognostic
//           fields, u cos(lat) and v cos(lat) from the stream function and
//           velocity potential (get_cos_lat_vector, spherical_harmonic.py:
//           1164-1178, clip=False), cos(lat) grad ln ps, and for moist
//           equations laplacian(ln ps) and cos(lat) grad q
//           (primitive_equations.py:168-225, 861-963).
// epilogue: the batched forward-transform output -> tendencies:
//           curl / divergence of the combined fluxes, kinetic energy (with the
//           moist geopotential correction folded in by the nodal kernel) and
//           orography Laplacians, div_sec_lat of the advective fluxes, the
//           humidity corrections and the top-wavenumber clip
//           (primitive_equations.py:1001-1059, 778-805, 1126-1194).
//
// Operators on a Real-layout (2M-1, L) field x (row r, degree l):
//   dlon(x)[r, l]  = jr * x[pr, l]   (jr = j or -j, pr = partner row; 0 at r=0)
//   cldl(x)[r, l]  = UP[r, l+1] x[r, l+1] + DN[r, l-1] x[r, l-1]
//   sdlc(x)[r, l]  = UP2[r, l+1] x[r, l+1] + DN2[r, l-1] x[r, l-1]
// with UP = (l+1) a_lm, DN = -l b_lm (cos_lat_d_dlat, :1063-1083) and
// UP2 = (l-1) a_lm, DN2 = -(l+2) b_lm (sec_lat_d_dlat_cos2, :1085-1092).
// dlon is antisymmetric; the adjoints of cldl/sdlc are the transposed
// two-point stencils.  One thread per (member, r, l) walks a chunk of levels.
#include <cstring>
#include <vector>

#include "dino_common.h"

struct dino_modal {
  int R, L, K, nt, hum;
  int hybrid;  // DINO_MODAL_HYBRID: prep also emits ln ps, no humidity extras
  double radius, gravity;
  // device constant block: il[L], ev[L], mask[L], UP, DN, UP2, DN2 (R*L each),
  // jr[R] (coefficient), pr[R] (partner row, stored as a number)
  double* c64;
  float* c32;
  // dense positions r * L + l with l >= leg_tile_start(m, L), m = (r + 1) / 2:
  // the part of a field the inverse Legendre kernels read, in row order
  int* tile_pos;
  int64_t n_tile_pos;
};

namespace dino {
namespace {

struct MOff {
  int il, ev, mask, up, dn, up2, dn2, jr, pr, zero, total;
};

__host__ __device__ inline MOff moff(int R, int L) {
  MOff o;
  o.il = 0;
  o.ev = L;
  o.mask = 2 * L;
  o.up = 3 * L;
  o.dn = o.up + R * L;
  o.up2 = o.dn + R * L;
  o.dn2 = o.up2 + R * L;
  o.jr = o.dn2 + R * L;
  o.pr = o.jr + R;
  o.zero = o.pr + R;  // R*L zeros: the field that threads below the triangle read
  o.total = o.zero + R * L;
  return o;
}

constexpr int MAXTR = 6;

struct PrepGroups {  // inverse-transform input (fields)
  int64_t z, d, t, tr[MAXTR], u, v, gx, gy, lp, qx, qy, ln, n;
};
struct StateGroups {  // state leaves in one buffer: vort, div, temp, lnps, tracers
  int64_t z, d, t, lnps, tr[MAXTR], n;
};
struct EpiGroups {  // forward-transform output (fields)
  int64_t cu, cv, ke, te, ut, vt, lsp, trn[MAXTR], tu[MAXTR], tv[MAXTR], curl, dc, n;
};

// hybrid: the sigma groups of a dry state plus ln ps itself (LN), which the
// hybrid nodal kernel needs for the column's surface pressure
inline PrepGroups prep_groups(int K, int nt, bool moist, int64_t E, bool hybrid = false) {
  PrepGroups g;
  std::memset(&g, 0, sizeof(g));
  int64_t o = 0;
  auto take = [&](int64_t lev) { int64_t r = o; o += lev * E; return r; };
  g.z = take(K); g.d = take(K); g.t = take(K);
  for (int i = 0; i < nt; ++i) g.tr[i] = take(K);
  g.u = take(K); g.v = take(K); g.gx = take(1); g.gy = take(1);
  if (moist) { g.lp = take(1); g.qx = take(K); g.qy = take(K); }
  if (hybrid) g.ln = take(1);
  g.n = o;
  return g;
}

inline StateGroups state_groups(int K, int nt, int64_t E) {
  StateGroups g;
  std::memset(&g, 0, sizeof(g));
  int64_t o = 0;
  auto take = [&](int64_t lev) { int64_t r = o; o += lev * E; return r; };
  g.z = take(K); g.d = take(K); g.t = take(K); g.lnps = take(1);
  for (int i = 0; i < nt; ++i) g.tr[i] = take(K);
  g.n = o;
  return g;
}

inline EpiGroups epi_groups(int K, int nt, bool moist, int64_t E) {
  EpiGroups g;
  std::memset(&g, 0, sizeof(g));
  int64_t o = 0;
  auto take = [&](int64_t lev) { int64_t r = o; o += lev * E; return r; };
  g.cu = take(K); g.cv = take(K); g.ke = take(K); g.te = take(K); g.ut = take(K);
  g.vt = take(K); g.lsp = take(1);
  for (int i = 0; i < nt; ++i) { g.trn[i] = take(K); g.tu[i] = take(K); g.tv[i] = take(K); }
  if (moist) { g.curl = take(K); g.dc = take(K); }
  g.n = o;
  return g;
}

struct MP {
  int R, L, K, nt, hum;  // hum < 0 also on hybrid levels (no humidity extras)
  int hybrid;
  int kc, nkc;  // levels per thread and level chunks: (logical) grid row = member * nkc + chunk
  int64_t RL;
  double a, g;
};

// Levels one thread walks: chunks of 4 levels give the launch ~9x more
// threads than whole columns at 37 levels, and 10-18% shorter kernels at
// TL255 x 16 members (more loads in flight; the per-position stencil
// coefficients are re-read per chunk, from cache).
#ifndef DINO_MODAL_LEVELS_PER_THREAD
#define DINO_MODAL_LEVELS_PER_THREAD 4
#endif

// Stencil of one thread's (r, l): the bidiagonal-in-l operators read the
// neighbours l - 1 and l + 1.  Out-of-range neighbours are loaded from a
// clamped index and discarded by a select, so no load sits behind a branch
// (a guarded load would wait on its own); the per-(r, l) coefficients are
// read once, since they are the same for every field and level.  Each
// operator adds its two terms in the same order as a sequential
// v = 0; v += first; v += second, with zero for a missing term.
template <typename T>
struct Stencil {
  int L, row, l, lm, lp, prow;
  bool hm, hp;
  T jr;
  T up_p, dn_m, up2_p, dn2_m;  // forward operators: coefficients at l + 1, l - 1
  T up_l, dn_l, up2_l, dn2_l;  // transposes: coefficients at l
  T il_p, il_m;                // inverse Laplacian at l + 1, l - 1
  __device__ Stencil(const T* c, MOff o, int R_, int L_, int r, int l_) : L(L_), l(l_) {
    (void)R_;
    row = r * L;
    lm = l > 0 ? l - 1 : 0;
    lp = l + 1 < L ? l + 1 : L - 1;
    // the neighbour below counts only inside the triangle (l - 1 >= m): the
    // degrees under it are zero, and a buffer written by a triangle-only
    // transform leaves some of them untouched
    hm = l - 1 >= ((r + 1) >> 1);
    hp = l + 1 < L;
    jr = c[o.jr + r];
    prow = (int)c[o.pr + r] * L;
    up_p = c[o.up + row + lp];
    dn_m = c[o.dn + row + lm];
    up2_p = c[o.up2 + row + lp];
    dn2_m = c[o.dn2 + row + lm];
    up_l = c[o.up + row + l];
    dn_l = c[o.dn + row + l];
    up2_l = c[o.up2 + row + l];
    dn2_l = c[o.dn2 + row + l];
    il_p = c[o.il + lp];
    il_m = c[o.il + lm];
  }
  __device__ T dlon(const T* x) const { return jr * x[prow + l]; }
  __device__ T cldl(const T* x) const {
    const T a = up_p * x[row + lp], b = dn_m * x[row + lm];
    return (hp ? a : (T)0) + (hm ? b : (T)0);
  }
  __device__ T sdlc(const T* x) const {
    const T a = up2_p * x[row + lp], b = dn2_m * x[row + lm];
    return (hp ? a : (T)0) + (hm ? b : (T)0);
  }
  // transposes: cldl^T(y)[l] = UP[l] y[l-1] + DN[l] y[l+1]
  __device__ T cldl_t(const T* y) const {
    const T a = up_l * y[row + lm], b = dn_l * y[row + lp];
    return (hm ? a : (T)0) + (hp ? b : (T)0);
  }
  __device__ T sdlc_t(const T* y) const {
    const T a = up2_l * y[row + lm], b = dn2_l * y[row + lp];
    return (hm ? a : (T)0) + (hp ? b : (T)0);
  }
  // cldl of (il * x): the inverse Laplacian applied at the stencil points
  __device__ T cldl_il(const T* x) const {
    const T a = up_p * il_p * x[row + lp], b = dn_m * il_m * x[row + lm];
    return (hp ? a : (T)0) + (hm ? b : (T)0);
  }
};

// ---------------------------------------------------------------- prep
// Only the inverse Legendre transform reads the output, from the first degree
// of its tiles on: threads walk the listed positions l >= leg_tile_start(m, L)
// only (full waves; the degrees below are not written).
template <typename T>
__global__ __launch_bounds__(256) void prep_kernel(const T* __restrict__ st, T* __restrict__ out,
                                                   const T* __restrict__ cst,
                                                   const int* __restrict__ tile_pos,
                                                   int64_t n_pos, StateGroups sg, PrepGroups pg,
                                                   MP p) {
  int bx, by;
  xcd_block(&bx, &by);
  const int64_t q = (int64_t)bx * blockDim.x + threadIdx.x;
  const int64_t e = by / p.nkc;
  const int kch = by % p.nkc;
  if (q >= n_pos) return;
  const int64_t pos = tile_pos[q];
  const int r = (int)(pos / p.L), l = (int)(pos % p.L);
  const MOff o = moff(p.R, p.L);
  Stencil<T> S(cst, o, p.R, p.L, r, l);
  const int K = p.K;
  const T inva = (T)(1.0 / p.a);
  const T il = cst[o.il + l];
  // Below the triangle (l < m) every field of the Real layout is zero, and so
  // is every output: those threads read a zero field in the constant block
  // instead of HBM and write zeros.
  const bool tri = l < ((r + 1) >> 1);
  const T* zb = cst + o.zero;
  auto F = [&](const T* base, int64_t off, int lev, int k) {
    return tri ? zb : base + (off + e * lev + k) * p.RL;
  };
  auto W = [&](T* base, int64_t off, int lev, int k) { return base + (off + e * lev + k) * p.RL; };
  const int k1 = (kch + 1) * p.kc < K ? (kch + 1) * p.kc : K;
  for (int k = kch * p.kc; k < k1; ++k) {
    const T* z = F(st, sg.z, K, k);
    const T* d = F(st, sg.d, K, k);
    T zv = z[pos], dv = d[pos];
    *(W(out, pg.z, K, k) + pos) = zv;
    *(W(out, pg.d, K, k) + pos) = dv;
    *(W(out, pg.t, K, k) + pos) = *(F(st, sg.t, K, k) + pos);
    // psi = il * zeta, chi = il * delta (il is diagonal in l)
    T dpsi = S.dlon(z) * il, dchi = S.dlon(d) * il;
    T cpsi = S.cldl_il(z), cchi = S.cldl_il(d);
    *(W(out, pg.u, K, k) + pos) = (dchi - cpsi) * inva;
    *(W(out, pg.v, K, k) + pos) = (cchi + dpsi) * inva;
    for (int i = 0; i < MAXTR; ++i) {
      if (i >= p.nt) break;
      *(W(out, pg.tr[i], K, k) + pos) = *(F(st, sg.tr[i], K, k) + pos);
    }
    if (p.hum >= 0) {
      const T* q = F(st, sg.tr[p.hum], K, k);
      *(W(out, pg.qx, K, k) + pos) = S.dlon(q) * inva;
      *(W(out, pg.qy, K, k) + pos) = S.cldl(q) * inva;
    }
  }
  if (kch != 0) return;  // the single-level fields, once per member
  const T* lp = F(st, sg.lnps, 1, 0);
  *(W(out, pg.gx, 1, 0) + pos) = S.dlon(lp) * inva;
  *(W(out, pg.gy, 1, 0) + pos) = S.cldl(lp) * inva;
  if (p.hum >= 0) *(W(out, pg.lp, 1, 0) + pos) = lp[pos] * cst[o.ev + l];
  if (p.hybrid) *(W(out, pg.ln, 1, 0) + pos) = lp[pos];
}

template <typename T>
__global__ __launch_bounds__(256) void prep_adjoint_kernel(const T* __restrict__ gout,
                                                           T* __restrict__ gst,
                                                           const T* __restrict__ cst,
                                                           StateGroups sg, PrepGroups pg, MP p) {
  int bx, by;
  xcd_block(&bx, &by);
  const int64_t pos = (int64_t)bx * blockDim.x + threadIdx.x;
  const int64_t e = by / p.nkc;
  const int kch = by % p.nkc;
  if (pos >= p.RL) return;
  const int r = (int)(pos / p.L), l = (int)(pos % p.L);
  const MOff o = moff(p.R, p.L);
  Stencil<T> S(cst, o, p.R, p.L, r, l);
  const int K = p.K;
  const T inva = (T)(1.0 / p.a);
  const T il = cst[o.il + l];
  // Below the triangle (l < m) every field of the Real layout is zero, and so
  // is every output: those threads read a zero field in the constant block
  // instead of HBM and write zeros.
  const bool tri = l < ((r + 1) >> 1);
  const T* zb = cst + o.zero;
  auto F = [&](const T* base, int64_t off, int lev, int k) {
    return tri ? zb : base + (off + e * lev + k) * p.RL;
  };
  auto W = [&](T* base, int64_t off, int lev, int k) { return base + (off + e * lev + k) * p.RL; };
  const int k1 = (kch + 1) * p.kc < K ? (kch + 1) * p.kc : K;
  for (int k = kch * p.kc; k < k1; ++k) {
    const T* ub = F(gout, pg.u, K, k);
    const T* vb = F(gout, pg.v, K, k);
    // U = (dlon(chi) - cldl(psi)) / a, V = (cldl(chi) + dlon(psi)) / a
    T psib = (-S.cldl_t(ub) - S.dlon(vb)) * inva;  // dlon^T = -dlon
    T chib = (-S.dlon(ub) + S.cldl_t(vb)) * inva;
    *(W(gst, sg.z, K, k) + pos) = *(F(gout, pg.z, K, k) + pos) + il * psib;
    *(W(gst, sg.d, K, k) + pos) = *(F(gout, pg.d, K, k) + pos) + il * chib;
    *(W(gst, sg.t, K, k) + pos) = *(F(gout, pg.t, K, k) + pos);
    for (int i = 0; i < MAXTR; ++i) {
      if (i >= p.nt) break;
      T v = *(F(gout, pg.tr[i], K, k) + pos);
      if (i == p.hum)
        v += (-S.dlon(F(gout, pg.qx, K, k)) + S.cldl_t(F(gout, pg.qy, K, k))) * inva;
      *(W(gst, sg.tr[i], K, k) + pos) = v;
    }
  }
  if (kch != 0) return;  // the single-level field, once per member
  T lb = (-S.dlon(F(gout, pg.gx, 1, 0)) + S.cldl_t(F(gout, pg.gy, 1, 0))) * inva;
  if (p.hum >= 0) lb += cst[o.ev + l] * *(F(gout, pg.lp, 1, 0) + pos);
  if (p.hybrid) lb += *(F(gout, pg.ln, 1, 0) + pos);
  *(W(gst, sg.lnps, 1, 0) + pos) = lb;
}

// ---------------------------------------------------------------- epilogue
template <typename T>
__global__ __launch_bounds__(256) void epilogue_kernel(const T* __restrict__ in,
                                                       const T* __restrict__ oro,
                                                       T* __restrict__ st,
                                                       const T* __restrict__ cst, EpiGroups eg,
                                                       StateGroups sg, MP p) {
  int bx, by;
  xcd_block(&bx, &by);
  const int64_t pos = (int64_t)bx * blockDim.x + threadIdx.x;
  const int64_t e = by / p.nkc;
  const int kch = by % p.nkc;
  if (pos >= p.RL) return;
  const int r = (int)(pos / p.L), l = (int)(pos % p.L);
  const MOff o = moff(p.R, p.L);
  Stencil<T> S(cst, o, p.R, p.L, r, l);
  const int K = p.K;
  const T inva = (T)(1.0 / p.a);
  const T ev = cst[o.ev + l], mask = cst[o.mask + l];
  const T orot = oro ? -(T)p.g * ev * oro[pos] : (T)0;
  // Below the triangle (l < m) every field of the Real layout is zero, and so
  // is every output: those threads read a zero field in the constant block
  // instead of HBM and write zeros.
  const bool tri = l < ((r + 1) >> 1);
  const T* zb = cst + o.zero;
  auto F = [&](const T* base, int64_t off, int lev, int k) {
    return tri ? zb : base + (off + e * lev + k) * p.RL;
  };
  auto W = [&](T* base, int64_t off, int lev, int k) { return base + (off + e * lev + k) * p.RL; };
  const bool moist = p.hum >= 0;
  const int k1 = (kch + 1) * p.kc < K ? (kch + 1) * p.kc : K;
  for (int k = kch * p.kc; k < k1; ++k) {
    const T* cu = F(in, eg.cu, K, k);
    const T* cv = F(in, eg.cv, K, k);
    // -curl_cos(cu, cv) = (sdlc(cu) - dlon(cv)) / a ; -div_cos = -(dlon(cu) + sdlc(cv)) / a
    T vort = (S.sdlc(cu) - S.dlon(cv)) * inva;
    T div = -(S.dlon(cu) + S.sdlc(cv)) * inva - ev * *(F(in, eg.ke, K, k) + pos) + orot;
    T temp = *(F(in, eg.te, K, k) + pos) -
             (S.dlon(F(in, eg.ut, K, k)) + S.sdlc(F(in, eg.vt, K, k))) * inva;
    if (moist) {
      vort += *(F(in, eg.curl, K, k) + pos);
      div -= *(F(in, eg.dc, K, k) + pos);
    }
    *(W(st, sg.z, K, k) + pos) = vort * mask;
    *(W(st, sg.d, K, k) + pos) = div * mask;
    *(W(st, sg.t, K, k) + pos) = temp * mask;
    for (int i = 0; i < MAXTR; ++i) {
      if (i >= p.nt) break;
      T v = *(F(in, eg.trn[i], K, k) + pos) -
            (S.dlon(F(in, eg.tu[i], K, k)) + S.sdlc(F(in, eg.tv[i], K, k))) * inva;
      *(W(st, sg.tr[i], K, k) + pos) = v * mask;
    }
  }
  if (kch == 0) *(W(st, sg.lnps, 1, 0) + pos) = *(F(in, eg.lsp, 1, 0) + pos) * mask;
}

// Output positions as for prep_kernel: the inverse Legendre transform of the
// adjoint reads it from the first degree of its tiles on.
template <typename T>
__global__ __launch_bounds__(256) void epilogue_adjoint_kernel(const T* __restrict__ gst,
                                                               T* __restrict__ gin,
                                                               const T* __restrict__ cst,
                                                               const int* __restrict__ tile_pos,
                                                               int64_t n_pos, EpiGroups eg,
                                                               StateGroups sg, MP p) {
  int bx, by;
  xcd_block(&bx, &by);
  const int64_t q = (int64_t)bx * blockDim.x + threadIdx.x;
  const int64_t e = by / p.nkc;
  const int kch = by % p.nkc;
  if (q >= n_pos) return;
  const int64_t pos = tile_pos[q];
  const int r = (int)(pos / p.L), l = (int)(pos % p.L);
  const MOff o = moff(p.R, p.L);
  Stencil<T> S(cst, o, p.R, p.L, r, l);
  const int K = p.K;
  const T inva = (T)(1.0 / p.a);
  const T ev = cst[o.ev + l];
  // Below the triangle (l < m) every field of the Real layout is zero, and so
  // is every output: those threads read a zero field in the constant block
  // instead of HBM and write zeros.
  const bool tri = l < ((r + 1) >> 1);
  const T* zb = cst + o.zero;
  auto F = [&](const T* base, int64_t off, int lev, int k) {
    return tri ? zb : base + (off + e * lev + k) * p.RL;
  };
  auto W = [&](T* base, int64_t off, int lev, int k) { return base + (off + e * lev + k) * p.RL; };
  // cotangents of the clipped outputs: y' = mask * ybar (mask is diagonal),
  // read through a masked view so stencil neighbours are masked too
  const T mk_l = cst[o.mask + l], mk_m = cst[o.mask + S.lm], mk_p = cst[o.mask + S.lp];
  auto MB = [&](const T* y) -> T { return y[S.row + l] * mk_l; };
  auto dlon_m = [&](const T* y) -> T { return S.jr * y[S.prow + l] * mk_l; };
  // sdlc^T of the masked cotangent (the neighbours' masks at l -+ 1)
  auto sdlc_t_m = [&](const T* y) -> T {
    const T a = S.up2_l * (y[S.row + S.lm] * mk_m), b = S.dn2_l * (y[S.row + S.lp] * mk_p);
    return (S.hm ? a : (T)0) + (S.hp ? b : (T)0);
  };
  const bool moist = p.hum >= 0;
  const int k1 = (kch + 1) * p.kc < K ? (kch + 1) * p.kc : K;
  for (int k = kch * p.kc; k < k1; ++k) {
    const T* zb = F(gst, sg.z, K, k);
    const T* db = F(gst, sg.d, K, k);
    const T* tb = F(gst, sg.t, K, k);
    const T zm = MB(zb), dm = MB(db), tm = MB(tb);
    // vort = (sdlc(cu) - dlon(cv)) / a ; div = -(dlon(cu) + sdlc(cv)) / a - ev ke ...
    T cub = (sdlc_t_m(zb) + dlon_m(db)) * inva;
    T cvb = (dlon_m(zb) - sdlc_t_m(db)) * inva;
    *(W(gin, eg.cu, K, k) + pos) = cub;
    *(W(gin, eg.cv, K, k) + pos) = cvb;
    *(W(gin, eg.ke, K, k) + pos) = -ev * dm;
    *(W(gin, eg.te, K, k) + pos) = tm;
    *(W(gin, eg.ut, K, k) + pos) = dlon_m(tb) * inva;
    *(W(gin, eg.vt, K, k) + pos) = -sdlc_t_m(tb) * inva;
    for (int i = 0; i < MAXTR; ++i) {
      if (i >= p.nt) break;
      const T* xb = F(gst, sg.tr[i], K, k);
      *(W(gin, eg.trn[i], K, k) + pos) = MB(xb);
      *(W(gin, eg.tu[i], K, k) + pos) = dlon_m(xb) * inva;
      *(W(gin, eg.tv[i], K, k) + pos) = -sdlc_t_m(xb) * inva;
    }
    if (moist) {
      *(W(gin, eg.curl, K, k) + pos) = zm;
      *(W(gin, eg.dc, K, k) + pos) = -dm;
    }
  }
  if (kch == 0) *(W(gin, eg.lsp, 1, 0) + pos) = MB(F(gst, sg.lnps, 1, 0));
}

template <typename T>
const T* consts(const dino_modal* m) {
  if constexpr (sizeof(T) == 8) return m->c64;
  else return m->c32;
}

MP make_mp(const dino_modal* m) {
  MP p;
  p.R = m->R;
  p.L = m->L;
  p.K = m->K;
  p.nt = m->nt;
  p.hum = m->hybrid ? -1 : m->hum;
  p.hybrid = m->hybrid;
  p.kc = DINO_MODAL_LEVELS_PER_THREAD < m->K ? DINO_MODAL_LEVELS_PER_THREAD : m->K;
  p.nkc = (m->K + p.kc - 1) / p.kc;
  p.RL = (int64_t)m->R * m->L;
  p.a = m->radius;
  p.g = m->gravity;
  return p;
}

template <typename T>
int run(int which, const dino_modal* m, const void* a, const void* b, void* out, int64_t E,
        hipStream_t s) {
  MP p = make_mp(m);
  // level chunks: as many as the grid's y extent allows for E members
  while ((int64_t)p.nkc * E > 65535) {
    p.kc *= 2;
    p.nkc = (m->K + p.kc - 1) / p.kc;
  }
  const bool moist = m->hum >= 0 && !m->hybrid, hyb = m->hybrid != 0;
  StateGroups sg = state_groups(m->K, m->nt, E);
  const unsigned gy = (unsigned)(E * p.nkc);
  dim3 grid((unsigned)((p.RL + 255) / 256), gy);
  dim3 tgrid((unsigned)((m->n_tile_pos + 255) / 256), gy);
  const T* c = consts<T>(m);
  switch (which) {
    case 0:
      prep_kernel<T><<<tgrid, 256, 0, s>>>((const T*)a, (T*)out, c, m->tile_pos, m->n_tile_pos,
                                           sg, prep_groups(m->K, m->nt, moist, E, hyb), p);
      break;
    case 1:
      prep_adjoint_kernel<T><<<grid, 256, 0, s>>>((const T*)a, (T*)out, c, sg,
                                                  prep_groups(m->K, m->nt, moist, E, hyb), p);
      break;
    case 2:
      epilogue_kernel<T><<<grid, 256, 0, s>>>((const T*)a, (const T*)b, (T*)out, c,
                                              epi_groups(m->K, m->nt, moist, E), sg, p);
      break;
    default:
      epilogue_adjoint_kernel<T><<<tgrid, 256, 0, s>>>((const T*)a, (T*)out, c, m->tile_pos,
                                                       m->n_tile_pos,
                                                       epi_groups(m->K, m->nt, moist, E), sg, p);
  }
  DINO_LAUNCH_CHECK("modal kernel");
  return DINO_OK;
}

}  // namespace
}  // namespace dino

using namespace dino;

extern "C" int dino_modal_create_ex(int R, int L, int K, int n_tracers, int humidity_index,
                                    double radius, double gravity, const double* inv_lap,
                                    const double* lap, const double* clip_mask, const double* up,
                                    const double* dn, const double* up2, const double* dn2,
                                    int flags, dino_modal** out) {
  DINO_REQUIRE(out, DINO_EINVAL, "dino_modal_create: out is NULL");
  DINO_REQUIRE((flags & ~DINO_MODAL_HYBRID) == 0, DINO_EINVAL, "unknown modal flags 0x%x",
               flags);
  *out = nullptr;
  DINO_REQUIRE(R >= 1 && R % 2 == 1 && L >= 2 && K >= 1 && (int64_t)R * L < (int64_t)1 << 31,
               DINO_EINVAL, "invalid modal shape (%d, %d) x %d levels", R, L, K);
  DINO_REQUIRE(n_tracers >= 0 && n_tracers <= MAXTR, DINO_EINVAL,
               "at most %d tracers are supported", MAXTR);
  DINO_REQUIRE(humidity_index < n_tracers, DINO_EINVAL, "humidity index out of range");
  DINO_REQUIRE(inv_lap && lap && clip_mask && up && dn && up2 && dn2, DINO_EINVAL,
               "NULL coefficient array");
  MOff o = moff(R, L);
  std::vector<double> h(o.total, 0.0);
  for (int l = 0; l < L; ++l) {
    h[o.il + l] = inv_lap[l];
    h[o.ev + l] = lap[l];
    h[o.mask + l] = clip_mask[l];
  }
  for (int i = 0; i < R * L; ++i) {
    h[o.up + i] = up[i];
    h[o.dn + i] = dn[i];
    h[o.up2 + i] = up2[i];
    h[o.dn2 + i] = dn2[i];
  }
  for (int r = 0; r < R; ++r) {
    int j = (r + 1) / 2;
    if (r == 0) {
      h[o.jr + r] = 0.0;
      h[o.pr + r] = 0.0;
    } else if (r & 1) {
      h[o.jr + r] = j;
      h[o.pr + r] = r + 1;
    } else {
      h[o.jr + r] = -j;
      h[o.pr + r] = r - 1;
    }
  }
  std::vector<float> hf(h.begin(), h.end());
  std::vector<int> tp;
  for (int r = 0; r < R; ++r)
    for (int l = leg_tile_start((r + 1) / 2, L); l < L; ++l) tp.push_back(r * L + l);
  dino_modal* m = new dino_modal();
  std::memset(m, 0, sizeof(*m));
  m->R = R;
  m->L = L;
  m->K = K;
  m->nt = n_tracers;
  m->hum = humidity_index;
  m->hybrid = (flags & DINO_MODAL_HYBRID) ? 1 : 0;
  m->radius = radius;
  m->gravity = gravity;
  hipError_t e = hipMalloc(&m->c64, sizeof(double) * o.total);
  if (e == hipSuccess) e = hipMalloc(&m->c32, sizeof(float) * o.total);
  if (e == hipSuccess) e = hipMemcpy(m->c64, h.data(), sizeof(double) * o.total, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->c32, hf.data(), sizeof(float) * o.total, hipMemcpyHostToDevice);
  m->n_tile_pos = (int64_t)tp.size();
  if (e == hipSuccess) e = hipMalloc(&m->tile_pos, sizeof(int) * tp.size());
  if (e == hipSuccess)
    e = hipMemcpy(m->tile_pos, tp.data(), sizeof(int) * tp.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (m->c64) (void)hipFree(m->c64);
    if (m->c32) (void)hipFree(m->c32);
    if (m->tile_pos) (void)hipFree(m->tile_pos);
    delete m;
    return hip_fail(e, "dino_modal_create");
  }
  *out = m;
  return DINO_OK;
}

extern "C" int dino_modal_create(int R, int L, int K, int n_tracers, int humidity_index,
                                 double radius, double gravity, const double* inv_lap,
                                 const double* lap, const double* clip_mask, const double* up,
                                 const double* dn, const double* up2, const double* dn2,
                                 dino_modal** out) {
  return dino_modal_create_ex(R, L, K, n_tracers, humidity_index, radius, gravity, inv_lap, lap,
                              clip_mask, up, dn, up2, dn2, 0, out);
}

extern "C" int dino_modal_destroy(dino_modal* m) {
  if (m) {
    if (m->c64) (void)hipFree(m->c64);
    if (m->c32) (void)hipFree(m->c32);
    if (m->tile_pos) (void)hipFree(m->tile_pos);
    delete m;
  }
  return DINO_OK;
}

extern "C" int64_t dino_modal_fields(const dino_modal* m, int64_t members, int which) {
  if (!m || members < 0) return -1;
  const bool moist = m->hum >= 0 && !m->hybrid;
  if (which == 0) return state_groups(m->K, m->nt, members).n;
  if (which == 1) return prep_groups(m->K, m->nt, moist, members, m->hybrid != 0).n;
  return epi_groups(m->K, m->nt, moist, members).n;
}

#define MODAL_ENTRY(NAME, WHICH)                                                            \
  extern "C" int NAME(const dino_modal* m, const void* a, const void* b, void* out,          \
                      int64_t members, int dtype, void* stream) {                           \
    DINO_REQUIRE(m && a && out, DINO_EINVAL, #NAME ": NULL argument");                      \
    DINO_REQUIRE(members >= 1 && members <= 65535, DINO_EINVAL, "members out of range");   \
    if (dtype == DINO_F64) return run<double>(WHICH, m, a, b, out, members, as_stream(stream)); \
    return run<float>(WHICH, m, a, b, out, members, as_stream(stream));                    \
  }

MODAL_ENTRY(dino_modal_prep, 0)
MODAL_ENTRY(dino_modal_prep_adjoint, 1)
MODAL_ENTRY(dino_modal_epilogue, 2)
MODAL_ENTRY(dino_modal_epilogue_adjoint, 3)
