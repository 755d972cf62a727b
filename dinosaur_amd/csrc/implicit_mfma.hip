// Semi-implicit solve and implicit terms on the FP64 matrix cores.
//
// Per column (member, modal row, degree l) of the flat state the solve is
//   z       = diag(1 / (1 - eta^2 lam_l d)) (W delta - eta lam_l (W G T + w lnps))
//   delta'' = V z,   T'' = T - eta (H V) z,   lnps'' = lnps - eta (dsigma^T V) z
// with A = G H + R T_ref dsigma^T = V diag(d) W (W = V^{-1}), w = W R T_ref
// (the same elimination as implicit.hip, primitive_equations.py:544-598,
// 1273-1330).  The implicit terms are
//   div = -lam_l (G T + R T_ref lnps),  temp = -H delta,  lnps = -dsigma . delta.
//
// All matrices are l-independent, so a batch of 16 columns is a pair of small
// FP64 GEMMs (v_mfma_f64_16x16x4f64).  A wave owns 16 columns:
//   stage 1: C_a = W [delta], C_b = (W G) [T] -- B operand lane (column j =
//            lane % 16, k-group kg = lane / 16) holds rows 4 s + kg;
//   stage 2: [V; H V; dsigma^T V] z -- the stage-1 accumulators are the B
//            operand as they stand: f64 accumulator register i of row tile it
//            holds row 16 it + 4 i + kg, which is row 4 s + kg of k-step
//            s = 4 it + i.  Only the k-steps s < KP / 4 carry nonzero rows of
//            z (rows >= K vanish), so stage 2 reduces over those alone.
// Matrix operands come from fragment-ordered tables built at create time,
// copied into LDS once per 8-wave block (one 8-byte LDS read per lane per
// MFMA, 512 contiguous bytes per wave).  Levels are
// padded to KP = 4 ceil(K / 4); the H V block starts at row KP, so the
// temperature fragments a lane loaded for stage 1 are the ones its stage-2
// output rows need, and the dsigma^T V row follows it at row KP + K (not
// 2 KP), so K = 37 takes 5 stage-2 row tiles instead of 6.
#include <cstring>
#include <functional>
#include <vector>

#include "dino_common.h"

namespace dino {
namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kWavesPerBlock = 8;  // waves sharing one LDS copy of the tables
// 16-column tiles each wave processes: the block's copy of the tables (tens
// of KiB from L2) then serves kRepeats x 8 tiles instead of 8
constexpr int kRepeats = 4;

__device__ __forceinline__ doublex4 mma64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// NS = KP / 4 reduction steps per stage-1 block and per stage-2 tile;
// T1 = row tiles of stage 1; T2MAX = the most row tiles stage 2 takes
// (KP + K + 1 <= 2 KP + 1 rows; the count for a given K is stage2_tiles).
template <int NS>
struct Shape {
  static constexpr int KP = 4 * NS;
  static constexpr int T1 = (KP + 15) / 16;
  static constexpr int T2MAX = (2 * KP + 1 + 15) / 16;
  static constexpr int TT = (KP + 1 + 15) / 16;  // terms: [H; dsigma] row tiles
};

__host__ __device__ inline int stage2_tiles(int K) {
  const int KP = 4 * ((K + 3) / 4);
  return (KP + K + 1 + 15) / 16;
}

template <typename T>
__device__ __forceinline__ double ldd(const T* p) {
  return (double)*p;
}

// TRANS = false: the solve above.  TRANS = true: its transpose (the
// vector-Jacobian product), which has the same two-stage shape:
//   z_bar  = D_l (V^T dd - eta ((H V)^T dT + (dsigma^T V)^T dp))
//   delta_bar = W^T z_bar, T_bar = dT - eta lam (W G)^T z_bar,
//   lnps_bar = dp - eta lam w . z_bar
// -- the tables are swapped by the host and eta / eta lam trade places.
// Below the triangle: for the Real layout (odd R, tri != 0) the column c of a
// member is (row r, degree l) with zonal wavenumber m = (r + 1) / 2, and
// fields vanish for l < m.  A wave whose 16 columns all lie there has zero
// input and zero output, so it reads nothing.
__device__ __forceinline__ bool below_triangle(int tri, bool cok, int r, int l) {
  const bool lane_tri = cok ? l < ((r + 1) >> 1) : true;
  return tri && __all(lane_tri);
}

// Column c (< cols = R L, so 32-bit) of a member: row r and degree l.
__device__ __forceinline__ void column_rl(int c, int L, int* r, int* l) {
  *r = c / L;
  *l = c - *r * L;
}

template <typename T>
__device__ __forceinline__ void zero_columns(T* div_out, T* temp_out, T* lnps_out, int64_t base,
                                             int64_t lcol, int64_t cols, int K, int kg, bool ok) {
  if (!ok) return;
  for (int k = kg; k < K; k += 4) {
    div_out[base + (int64_t)k * cols] = (T)0;
    temp_out[base + (int64_t)k * cols] = (T)0;
  }
  if (kg == 0) lnps_out[lcol] = (T)0;
}

// In place is allowed (outputs equal to inputs): a wave reads every level of
// its 16 columns before it writes any, and no other wave touches them.
template <typename T, int NS, bool TRANS>
__global__ __launch_bounds__(64 * kWavesPerBlock, (sizeof(T) == 4 && NS <= 10) ? 4 : 1)
void solve_mfma_kernel(
    const T* div, const T* temp, const T* lnps, T* div_out, T* temp_out, T* lnps_out,
    const double* __restrict__ t1a_g, const double* __restrict__ t1b_g,
    const double* __restrict__ t2_g, const double* __restrict__ wvec,
    const double* __restrict__ dvec, const double* __restrict__ lam, int K, int L,
    int64_t cols, int64_t members, double eta, int tri, int inplace) {
  constexpr bool ACCUM = false;
  using S = Shape<NS>;
  // the block's waves share the operand tables through LDS (one L2 read per
  // block instead of one per wave and MFMA)
  constexpr int N1 = S::T1 * NS * 64, N2 = S::T2MAX * NS * 64;
  // (and the per-row vectors of the z scaling, read per tile)
  __shared__ double t1a[N1], t1b[N1], t2[N2], wv[16 * S::T1], dv[16 * S::T1];
  for (int i = threadIdx.x; i < N1; i += blockDim.x) {
    t1a[i] = t1a_g[i];
    t1b[i] = t1b_g[i];
  }
  const int T2 = stage2_tiles(K);
  for (int i = threadIdx.x; i < T2 * NS * 64; i += blockDim.x) t2[i] = t2_g[i];
  for (int i = threadIdx.x; i < 16 * S::T1; i += blockDim.x) {
    wv[i] = wvec[i];
    dv[i] = dvec[i];
  }
  __syncthreads();
  const int tiles_per_member = (int)((cols + 15) / 16);
  // one 16-column tile (a return ends this tile only)
  auto tile = [&](int wave, int lane) {
  const int b = wave / tiles_per_member;
  const int c = (wave - b * tiles_per_member) * 16 + (lane & 15);
  const int kg = lane >> 4;
  const bool cok = c < cols;
  const int cc = cok ? c : 0;
  int r, l;
  column_rl(cc, L, &r, &l);
  const int64_t base = (int64_t)b * K * cols + cc;
  if (below_triangle(tri, cok, r, l)) {
    // every column of the wave has l < m: zero in, zero out
    if (!inplace && !ACCUM) zero_columns<T>(div_out, temp_out, lnps_out, base,
                                            (int64_t)b * cols + c, cols, K, kg, cok);
    return;
  }
  const double lam_l = lam[l];
  const double lp = ldd(lnps + (int64_t)b * cols + cc);
  // stage-1 B fragments: delta and T rows 4 s + kg (zero beyond K)
  double xd[NS], xt[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int k = 4 * s + kg;
    const bool ok = k < K;
    const double dv = ldd(div + base + (int64_t)(ok ? k : 0) * cols);
    const double tv = ldd(temp + base + (int64_t)(ok ? k : 0) * cols);
    xd[s] = ok ? dv : 0.0;
    xt[s] = ok ? tv : 0.0;
  }
  // all row tiles of stage 1 at once: 2 T1 independent accumulator chains
  // keep the matrix cores busy (one tile at a time, 2 chains, is slower)
  doublex4 ca[S::T1], cb[S::T1];
#pragma unroll
  for (int it = 0; it < S::T1; ++it) {
    ca[it] = doublex4{0.0, 0.0, 0.0, 0.0};
    cb[it] = ca[it];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      ca[it] = mma64(t1a[(it * NS + s) * 64 + lane], xd[s], ca[it]);
      cb[it] = mma64(t1b[(it * NS + s) * 64 + lane], xt[s], cb[it]);
    }
  }
  // z = (C_a - eta lam (C_b + w lnps)) / (1 - eta^2 lam d), rows >= K vanish;
  // stage 2 reads the k-steps s < NS (rows < KP) only
  double z[NS];
#pragma unroll
  for (int it = 0; it < S::T1; ++it) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (4 * it + i >= NS) continue;
      const int row = 16 * it + 4 * i + kg;
      const double c1 = TRANS ? eta : eta * lam_l;
      const double rhs = ca[it][i] - c1 * (cb[it][i] + wv[row] * lp);
      z[4 * it + i] = rhs / (1.0 - eta * eta * lam_l * dv[row]);
    }
  }
  // stage 2, one output row tile at a time: rows [0, KP) delta, [KP, KP + K)
  // temperature, KP + K ln ps
  const double c2 = TRANS ? eta * lam_l : eta;
  const int lrow = S::KP + K;
  double sacc = 0.0;
  auto emit = [&](int gt, const doublex4& acc) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 16 * gt + 4 * i + kg;
      if (row < S::KP) {
        if (cok && row < K) div_out[base + (int64_t)row * cols] = (T)acc[i];
      } else if (row < lrow) {
        const int q = row - S::KP;  // = 4 s + kg with s = 4 gt + i - NS
        const int s = 4 * gt + i - NS;
        if (cok && s >= 0 && s < NS)
          temp_out[base + (int64_t)q * cols] = (T)(xt[s] - c2 * acc[i]);
      } else if (row == lrow) {
        sacc = acc[i];
      }
    }
  };
  // output row tiles in pairs: two independent accumulator chains back to
  // back (all of them at once would hold every table fragment in registers)
#pragma unroll
  for (int g0 = 0; g0 < S::T2MAX; g0 += 2) {
    if (g0 >= T2) continue;  // uniform: K = KP takes the last tile
    const bool two = g0 + 1 < T2;
    doublex4 a0 = {0.0, 0.0, 0.0, 0.0}, a1 = a0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      a0 = mma64(t2[(g0 * NS + s) * 64 + lane], z[s], a0);
      if (two) a1 = mma64(t2[((g0 + 1) * NS + s) * 64 + lane], z[s], a1);
    }
    emit(g0, a0);
    if (two) emit(g0 + 1, a1);
  }
  // the lnps row lives in one k-group
  if (kg == (lrow & 3) && cok) lnps_out[(int64_t)b * cols + cc] = (T)(lp - c2 * sacc);
  };
  const int first = (int)blockIdx.x * kRepeats * kWavesPerBlock + (int)(threadIdx.x >> 6);
  const int n_waves = tiles_per_member * (int)members;
#pragma unroll 1
  for (int rep = 0; rep < kRepeats; ++rep) {
    const int wave = first + rep * kWavesPerBlock;
    if (wave >= n_waves) break;
    // an opaque copy of the lane index keeps the table reads inside the loop
    // (hoisted out of it they would hold ~220 registers)
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    tile(wave, lane);
  }
}

// TRANS = false: the implicit terms.  TRANS = true: their transpose
//   delta_bar = -(H^T dT + dsigma dp),  T_bar = -lam G^T dd,
//   lnps_bar = -lam (R T_ref) . dd
// with tg = H^T, th = [G^T; (R T_ref)^T], v1 = dsigma (host-side swap).
// ACC: add the result to the outputs instead of overwriting them.  In place
// is allowed, as for the solve.
template <typename T, int NS, bool TRANS, bool ACC>
__global__ __launch_bounds__(64 * kWavesPerBlock) void terms_mfma_kernel(
    const T* div, const T* temp, const T* lnps, T* div_out, T* temp_out, T* lnps_out,
    const double* __restrict__ tg_g, const double* __restrict__ th_g,
    const double* __restrict__ rtref, const double* __restrict__ lam, int K, int L,
    int64_t cols, int64_t members, int tri, int inplace) {
  constexpr bool ACCUM = ACC;  // adding the zero result changes nothing
  using S = Shape<NS>;
  constexpr int NG = S::T1 * NS * 64, NH = S::TT * NS * 64;
  __shared__ double tg[NG], th[NH], rt[16 * S::T1];
  for (int i = threadIdx.x; i < NG; i += blockDim.x) tg[i] = tg_g[i];
  for (int i = threadIdx.x; i < NH; i += blockDim.x) th[i] = th_g[i];
  for (int i = threadIdx.x; i < 16 * S::T1; i += blockDim.x) rt[i] = rtref[i];
  __syncthreads();
  const int tiles_per_member = (int)((cols + 15) / 16);
  // one 16-column tile (a return ends this tile only)
  auto tile = [&](int wave, int lane) {
  const int b = wave / tiles_per_member;
  const int c = (wave - b * tiles_per_member) * 16 + (lane & 15);
  const int kg = lane >> 4;
  const bool cok = c < cols;
  const int cc = cok ? c : 0;
  int r, l;
  column_rl(cc, L, &r, &l);
  const int64_t base = (int64_t)b * K * cols + cc;
  if (below_triangle(tri, cok, r, l)) {
    // every column of the wave has l < m: zero in, zero out
    if (!inplace && !ACCUM) zero_columns<T>(div_out, temp_out, lnps_out, base,
                                            (int64_t)b * cols + c, cols, K, kg, cok);
    return;
  }
  const double lam_l = lam[l];
  const double lp = ldd(lnps + (int64_t)b * cols + cc);
  double xd[NS], xt[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int k = 4 * s + kg;
    const bool ok = k < K;
    const double dv = ldd(div + base + (int64_t)(ok ? k : 0) * cols);
    const double tv = ldd(temp + base + (int64_t)(ok ? k : 0) * cols);
    xd[s] = ok ? dv : 0.0;
    xt[s] = ok ? tv : 0.0;
  }
  // divergence tendency -lam (G T + R T_ref lnps) (row tiles it < T1) and
  // temperature and log-pressure tendencies -[H; dsigma] delta (it < TT):
  // the two accumulator chains of a row tile index run back to back.  When
  // accumulating into the outputs (loads and stores of outputs that may
  // alias), the divergence tiles all come first, then the others (faster
  // there).
  const double s1 = TRANS ? 1.0 : lam_l;
  const double s2 = TRANS ? lam_l : 1.0;
  double sacc = 0.0;
  constexpr int TM = S::T1 > S::TT ? S::T1 : S::TT;
  constexpr int PH = ACC ? 2 : 1;
#pragma unroll
  for (int ph = 0; ph < PH; ++ph) {
#pragma unroll
  for (int it = 0; it < TM; ++it) {
    const bool dg = it < S::T1 && (!ACC || ph == 0);
    const bool dh = it < S::TT && (!ACC || ph == 1);
    if (!dg && !dh) continue;
    doublex4 ag = {0.0, 0.0, 0.0, 0.0}, ah = ag;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (dg) ag = mma64(tg[(it * NS + s) * 64 + lane], xt[s], ag);
      if (dh) ah = mma64(th[(it * NS + s) * 64 + lane], xd[s], ah);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 16 * it + 4 * i + kg;
      if (dg && cok && row < K) {
        T* o = div_out + base + (int64_t)row * cols;
        const T v = (T)(-s1 * (ag[i] + rt[row] * lp));
        *o = ACC ? (T)(*o + v) : v;
      }
      if (dh) {
        if (row < K) {
          if (cok) {
            T* o = temp_out + base + (int64_t)row * cols;
            const T v = (T)(-s2 * ah[i]);
            *o = ACC ? (T)(*o + v) : v;
          }
        } else if (row == S::KP) {
          sacc = ah[i];
        }
      }
    }
  }
  }
  if (kg == (S::KP & 3) && cok) {
    T* o = lnps_out + (int64_t)b * cols + cc;
    const T v = (T)(-s2 * sacc);
    *o = ACC ? (T)(*o + v) : v;
  }
  };
  const int first = (int)blockIdx.x * kRepeats * kWavesPerBlock + (int)(threadIdx.x >> 6);
  const int n_waves = tiles_per_member * (int)members;
#pragma unroll 1
  for (int rep = 0; rep < kRepeats; ++rep) {
    const int wave = first + rep * kWavesPerBlock;
    if (wave >= n_waves) break;
    // an opaque copy of the lane index keeps the table reads inside the loop
    // (hoisted out of it they would hold ~220 registers)
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    tile(wave, lane);
  }
}

}  // namespace

// Host-side construction of the fragment tables.  A table for a (rows x kdim)
// matrix M in tiles of 16 rows and reduction steps of 4:
//   tab[(tile * steps + s) * 64 + lane] = M[16 tile + lane % 16][4 s + lane / 16].
static void fragment_table(std::vector<double>& tab, int tiles, int steps,
                           const std::function<double(int, int)>& M) {
  tab.assign((size_t)tiles * steps * 64, 0.0);
  for (int t = 0; t < tiles; ++t)
    for (int s = 0; s < steps; ++s)
      for (int lane = 0; lane < 64; ++lane)
        tab[((size_t)t * steps + s) * 64 + lane] = M(16 * t + lane % 16, 4 * s + lane / 16);
}

int implicit_mfma_steps(int K) { return (K + 3) / 4; }

// Builds the tables for NS = ceil(K / 4).  Inputs are row-major K x K FP64.
int implicit_mfma_tables(int K, const double* G, const double* H, const double* tref,
                         const double* dsigma, const double* V, const double* Vinv,
                         const double* d, double r_gas, std::vector<double>* out) {
  const int NS = implicit_mfma_steps(K), KP = 4 * NS;
  const int T1 = (KP + 15) / 16, T2 = stage2_tiles(K);
  const int TT = (KP + 1 + 15) / 16;
  auto at = [&](const double* A, int r, int c) { return A[(size_t)r * K + c]; };
  std::vector<double> WG((size_t)K * K, 0.0), HV((size_t)K * K, 0.0), w(K, 0.0), sV(K, 0.0);
  if (V && Vinv) {
    for (int r = 0; r < K; ++r)
      for (int c = 0; c < K; ++c) {
        double a = 0.0, h = 0.0;
        for (int k = 0; k < K; ++k) {
          a += at(Vinv, r, k) * at(G, k, c);
          h += at(H, r, k) * at(V, k, c);
        }
        WG[(size_t)r * K + c] = a;
        HV[(size_t)r * K + c] = h;
      }
    for (int r = 0; r < K; ++r) {
      double a = 0.0, s = 0.0;
      for (int k = 0; k < K; ++k) {
        a += at(Vinv, r, k) * r_gas * tref[k];
        s += dsigma[k] * at(V, k, r);
      }
      w[r] = a;
      sV[r] = s;
    }
  }
  auto in = [&](int r, int c) { return r < K && c < K; };
  auto vec16 = [&](std::vector<double>& v, const std::function<double(int)>& f) {
    v.assign(16 * T1, 0.0);
    for (int r = 0; r < K; ++r) v[r] = f(r);
  };
  const bool modes = V && Vinv;
  // forward: [0] t1a = W, [1] t1b = W G, [2] t2 = [V; H V; dsigma^T V], [3] w,
  // [4] d, [5] tg = G, [6] th = [H; dsigma^T], [7] R T_ref
  fragment_table(out[0], T1, NS, [&](int r, int c) { return modes && in(r, c) ? at(Vinv, r, c) : 0.0; });
  fragment_table(out[1], T1, NS, [&](int r, int c) { return in(r, c) ? WG[(size_t)r * K + c] : 0.0; });
  fragment_table(out[2], T2, NS, [&](int r, int c) {
    if (!modes || c >= K) return 0.0;
    if (r < KP) return r < K ? at(V, r, c) : 0.0;
    if (r < KP + K) return HV[(size_t)(r - KP) * K + c];
    return r == KP + K ? sV[c] : 0.0;
  });
  vec16(out[3], [&](int r) { return w[r]; });
  vec16(out[4], [&](int r) { return d ? d[r] : 0.0; });
  fragment_table(out[5], T1, NS, [&](int r, int c) { return in(r, c) ? at(G, r, c) : 0.0; });
  fragment_table(out[6], TT, NS, [&](int r, int c) {
    if (c >= K) return 0.0;
    if (r < K) return at(H, r, c);
    return r == KP ? dsigma[c] : 0.0;
  });
  vec16(out[7], [&](int r) { return r_gas * tref[r]; });
  // transposes: [8] V^T, [9] (H V)^T, [10] [W^T; (W G)^T; w^T], [11] dsigma^T V,
  // [12] H^T, [13] [G^T; (R T_ref)^T], [14] dsigma
  fragment_table(out[8], T1, NS, [&](int r, int c) { return modes && in(r, c) ? at(V, c, r) : 0.0; });
  fragment_table(out[9], T1, NS, [&](int r, int c) { return in(r, c) ? HV[(size_t)c * K + r] : 0.0; });
  fragment_table(out[10], T2, NS, [&](int r, int c) {
    if (!modes || c >= K) return 0.0;
    if (r < KP) return r < K ? at(Vinv, c, r) : 0.0;
    if (r < KP + K) return WG[(size_t)c * K + (r - KP)];
    return r == KP + K ? w[c] : 0.0;
  });
  vec16(out[11], [&](int r) { return sV[r]; });
  fragment_table(out[12], T1, NS, [&](int r, int c) { return in(r, c) ? at(H, c, r) : 0.0; });
  fragment_table(out[13], TT, NS, [&](int r, int c) {
    if (c >= K) return 0.0;
    if (r < K) return at(G, c, r);
    return r == KP ? r_gas * tref[c] : 0.0;
  });
  vec16(out[14], [&](int r) { return dsigma[r]; });
  return DINO_OK;
}

template <typename T, int NS>
static int launch_ns(bool solve, bool adjoint, bool acc, const dino_implicit* f,
                     const void* div, const void* temp, const void* lnps, void* d_o, void* t_o,
                     void* l_o, int64_t members, hipStream_t s) {
  const int64_t cols = (int64_t)f->R * f->L;
  const int64_t waves = (cols + 15) / 16 * members;
  if (waves == 0) return DINO_OK;
  // column tiles and columns are indexed with 32-bit integers in the kernels
  DINO_REQUIRE(waves + (int64_t)kWavesPerBlock * kRepeats < ((int64_t)1 << 31) &&
                   cols < ((int64_t)1 << 31),
               DINO_EINVAL, "implicit solve: too many columns for one launch");
  const int64_t per_block = (int64_t)kWavesPerBlock * kRepeats;
  const unsigned blocks = (unsigned)((waves + per_block - 1) / per_block);
  const unsigned threads = 64 * kWavesPerBlock;
  const T* x0 = (const T*)div;
  const T* x1 = (const T*)temp;
  const T* x2 = (const T*)lnps;
  T* y0 = (T*)d_o;
  T* y1 = (T*)t_o;
  T* y2 = (T*)l_o;
  double* const* t = f->mtab;
  const int tri = (f->R % 2 == 1) ? 1 : 0;  // Real layout (odd row count)
  const int inplace = (d_o == div && t_o == temp && l_o == lnps) ? 1 : 0;
  if (solve && !adjoint) {
    solve_mfma_kernel<T, NS, false><<<blocks, threads, 0, s>>>(
        x0, x1, x2, y0, y1, y2, t[0], t[1], t[2], t[3], t[4], f->lam, f->K, f->L, cols,
        members, f->eta, tri, inplace);
  } else if (solve) {
    solve_mfma_kernel<T, NS, true><<<blocks, threads, 0, s>>>(
        x0, x1, x2, y0, y1, y2, t[8], t[9], t[10], t[11], t[4], f->lam, f->K, f->L, cols,
        members, f->eta, tri, inplace);
  } else if (!adjoint) {
    terms_mfma_kernel<T, NS, false, false><<<blocks, threads, 0, s>>>(
        x0, x1, x2, y0, y1, y2, t[5], t[6], t[7], f->lam, f->K, f->L, cols, members, tri,
        inplace);
  } else if (!acc) {
    terms_mfma_kernel<T, NS, true, false><<<blocks, threads, 0, s>>>(
        x0, x1, x2, y0, y1, y2, t[12], t[13], t[14], f->lam, f->K, f->L, cols, members, tri,
        inplace);
  } else {
    terms_mfma_kernel<T, NS, true, true><<<blocks, threads, 0, s>>>(
        x0, x1, x2, y0, y1, y2, t[12], t[13], t[14], f->lam, f->K, f->L, cols, members, tri,
        inplace);
  }
  DINO_LAUNCH_CHECK("implicit mfma kernel");
  return DINO_OK;
}

template <typename T>
int implicit_mfma_dispatch(bool solve, bool adjoint, bool acc, const dino_implicit* f,
                           const void* div, const void* temp, const void* lnps, void* d_o,
                           void* t_o, void* l_o, int64_t members, hipStream_t s) {
  switch (implicit_mfma_steps(f->K)) {
#define NS_CASE(n) \
  case n:          \
    return launch_ns<T, n>(solve, adjoint, acc, f, div, temp, lnps, d_o, t_o, l_o, members, s);
    NS_CASE(1) NS_CASE(2) NS_CASE(3) NS_CASE(4) NS_CASE(5) NS_CASE(6) NS_CASE(7) NS_CASE(8)
    NS_CASE(9) NS_CASE(10) NS_CASE(11) NS_CASE(12) NS_CASE(13) NS_CASE(14) NS_CASE(15)
    NS_CASE(16)
#undef NS_CASE
    default:
      set_error("implicit solve supports up to 64 levels, got %d", f->K);
      return DINO_EINVAL;
  }
}

template int implicit_mfma_dispatch<float>(bool, bool, bool, const dino_implicit*, const void*,
                                           const void*, const void*, void*, void*, void*,
                                           int64_t, hipStream_t);
template int implicit_mfma_dispatch<double>(bool, bool, bool, const dino_implicit*,
                                            const void*, const void*, const void*, void*, void*,
                                            void*, int64_t, hipStream_t);

}  // namespace dino
