// Legendre GEMMs on the CDNA4 matrix cores.
//
// Same contraction as legendre.hip (per zonal wavenumber m, triangular slab
// l >= m, equatorial fold), tiled for 64-wide wavefronts and MFMA 16x16 tiles:
//
// inverse (modal -> fourier): C[row, jj] = sum_l A[row, l] P[l, jj] for the
//   even-(l-m) and odd-(l-m) degrees separately (accE / accO); the epilogue
//   writes E + O at the northern node and E - O at its southern mirror.
//   Workgroup tile: 128 rows x TC latitudes, 4 waves as 2 x 2.
// forward (fourier -> modal): D[row, l] = sum_jj F[row, jj] P[l, jj] with
//   F = S (= f_up + f_dn) for even (l-m) and A (= f_up - f_dn) for odd.
//   Workgroup tile: 128 rows x 128 degrees (64 even + 64 odd).
//
// Operands are staged in LDS de-interleaved by parity.  Arithmetic:
//   FP32  : v_mfma_f32_16x16x4_f32 (exact FP32 products, FP32 accumulation)
//   split : hi + lo BF16 halves of both operands, three
//           v_mfma_f32_16x16x32 products (hi*hi + hi*lo + lo*hi) accumulated
//           in FP32.
#include "dino_common.h"

namespace dino {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;
constexpr int TR = 128;  // rows per workgroup tile

__device__ __forceinline__ void row_of(int m, int q, int* b, int* r) {
  if (m == 0) {
    *b = q;
    *r = 0;
  } else {
    *b = q >> 1;
    *r = 2 * m - 1 + (q & 1);
  }
}

// ------------------------------------------------------------ operand policies
struct PolicyF32 {
  static constexpr int KSTEP = 4;
  // A-operand: lane holds A[row0 + lane%16][k0 + lane/16], from LDS [k][row]
  template <int PA, int PB>
  __device__ static floatx4 mma(const float (*As)[PA], int arow, const float (*Bs)[PB],
                                int bcol, int k0, int lane, floatx4 acc) {
    float a = As[k0 + lane / 16][arow + lane % 16];
    float b = Bs[k0 + lane / 16][bcol + lane % 16];
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
};

struct PolicySplit {
  static constexpr int KSTEP = 32;
  template <int PA, int PB>
  __device__ static floatx4 mma(const float (*As)[PA], int arow, const float (*Bs)[PB],
                                int bcol, int k0, int lane, floatx4 acc) {
    const int kb = k0 + 8 * (lane / 16);
    const int ra = arow + lane % 16, cb = bcol + lane % 16;
    bf16x8 ah, al, bh, bl;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float a = As[kb + i][ra], b = Bs[kb + i][cb];
      __bf16 h = (__bf16)a, g = (__bf16)b;
      ah[i] = h;
      al[i] = (__bf16)(a - (float)h);
      bh[i] = g;
      bl[i] = (__bf16)(b - (float)g);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
  }
};

// ------------------------------------------------------------ inverse
// KH degrees per parity per chunk (chunk = 2 KH consecutive degrees).  The
// next chunk is fetched into registers while the current one is multiplied.
template <class Pol, int TC, int KH>
__global__ __launch_bounds__(kThreads) void leg_inv_mfma_kernel(
    const float* __restrict__ modal, float* __restrict__ fourier, const float* __restrict__ ptab,
    const int64_t* __restrict__ m_off, int B, int L, int J, int Jh, int Jhp, int Rrows,
    int mpair) {
  constexpr int WR = TR / 2, WC = TC / 2, MT = WR / 16, NT = WC / 16;
  constexpr int KC = 2 * KH;
  constexpr int PA = TR + 4, PB = TC + 4;
  constexpr int A_PER = TR * KC / kThreads, P_PER = KC * TC / kThreads;
  static_assert(TR * KC % kThreads == 0 && KC * TC % kThreads == 0, "tile / thread mismatch");
  const int m = m_of_block(blockIdx.z, gridDim.z, mpair);
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.y * TR;
  if (row0 >= nrows) return;
  const int jj0 = blockIdx.x * TC;
  __shared__ float Ae[KH][PA], Ao[KH][PA];
  __shared__ float Pe[KH][PB], Po[KH][PB];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int wr = wave / 2, wc = wave % 2;
  floatx4 accE[MT][NT], accO[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int c = 0; c < NT; ++c) accE[a][c] = accO[a][c] = floatx4{0, 0, 0, 0};
  const float* pm = ptab + m_off[m];
  // The rows a thread loads are q0 + i * (kThreads / KC): a fixed stride of
  // rows, i.e. (m == 0 ? 1 : 1/2) * stride fields in memory.
  constexpr int QS = kThreads / KC;
  static_assert(kThreads % KC == 0 && QS % 2 == 0, "row stride must be even");
  const int q0 = tid / KC, kq = tid % KC;
  const float* abase = nullptr;
  {
    int b, r;
    row_of(m, row0 + q0, &b, &r);
    abase = modal + ((int64_t)b * Rrows + r) * L;
  }
  const int64_t astride = (int64_t)(m == 0 ? QS : QS / 2) * Rrows * L;
  float ra[A_PER], rp[P_PER];
  auto fetch = [&](int l0) {
    const int l = l0 + kq;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      bool ok = (row0 + q0 + i * QS < nrows) && l < L;
      ra[i] = ok ? abase[i * astride + l] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < P_PER; ++i) {
      int idx = tid + i * kThreads;
      int k = idx / TC, c = idx % TC;
      int l = l0 + k, jj = jj0 + c;
      rp[i] = (l < L && jj < Jhp) ? pm[(int64_t)(l - m) * Jhp + jj] : 0.f;
    }
  };
  fetch(m);
  for (int l0 = m; l0 < L; l0 += KC) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int idx = tid + i * kThreads;
      int q = idx / KC, k = idx % KC;
      if (k & 1) Ao[k >> 1][q] = ra[i];
      else Ae[k >> 1][q] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < P_PER; ++i) {
      int idx = tid + i * kThreads;
      int k = idx / TC, c = idx % TC;
      if (k & 1) Po[k >> 1][c] = rp[i];
      else Pe[k >> 1][c] = rp[i];
    }
    __syncthreads();
    if (l0 + KC < L) fetch(l0 + KC);
#pragma unroll
    for (int ks = 0; ks < KH; ks += Pol::KSTEP) {
#pragma unroll
      for (int a = 0; a < MT; ++a) {
#pragma unroll
        for (int c = 0; c < NT; ++c) {
          int ar = wr * WR + a * 16, bc = wc * WC + c * 16;
          accE[a][c] = Pol::template mma<PA, PB>(Ae, ar, Pe, bc, ks, lane, accE[a][c]);
          accO[a][c] = Pol::template mma<PA, PB>(Ao, ar, Po, bc, ks, lane, accO[a][c]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < MT; ++a) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int row = row0 + wr * WR + a * 16 + 4 * (lane / 16) + i;
      if (row >= nrows) continue;
      int b, r;
      row_of(m, row, &b, &r);
      float* out = fourier + ((int64_t)b * Rrows + r) * J;
#pragma unroll
      for (int c = 0; c < NT; ++c) {
        int jj = jj0 + wc * WC + c * 16 + lane % 16;
        if (jj >= Jh) continue;
        int jup = J - Jh + jj, jdn = Jh - 1 - jj;
        float e = accE[a][c][i], o = accO[a][c][i];
        out[jup] = e + o;
        if (jdn != jup) out[jdn] = e - o;
      }
    }
  }
}

// ------------------------------------------------------------ forward
// Degree tiles of 128 cover [0, L).  Within a tile, even columns c pair with
// S and odd with A when (lt0 - m) is even (roles swap otherwise); P is zero
// for l < m, and tiles entirely below the diagonal only write zeros.
template <class Pol, int KC>
__global__ __launch_bounds__(kThreads) void leg_fwd_mfma_kernel(
    const float* __restrict__ fourier, float* __restrict__ modal, const float* __restrict__ ptab,
    const int64_t* __restrict__ m_off, int B, int L, int J, int Jh, int Jhp, int Rrows,
    int mpair) {
  constexpr int TLH = 64;  // degrees per parity in the tile
  constexpr int WR = TR / 2, WC = TLH / 2, MT = WR / 16, NT = WC / 16;
  constexpr int PA = TR + 4, PB = TLH + 4;
  constexpr int S_PER = TR * KC / kThreads, P_PER = 2 * TLH * KC / kThreads;
  static_assert(TR * KC % kThreads == 0 && 2 * TLH * KC % kThreads == 0, "tile mismatch");
  const int m = m_of_block(blockIdx.z, gridDim.z, mpair);
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.y * TR;
  if (row0 >= nrows) return;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int lt0 = blockIdx.x * (2 * TLH);
  if (lt0 >= L) return;
  if (lt0 + 2 * TLH <= m) {
    for (int idx = tid; idx < TR * 2 * TLH; idx += kThreads) {
      int q = idx / (2 * TLH), c = idx % (2 * TLH);
      int row = row0 + q, l = lt0 + c;
      if (row < nrows && l < L) {
        int b, r;
        row_of(m, row, &b, &r);
        modal[((int64_t)b * Rrows + r) * L + l] = 0.f;
      }
    }
    return;
  }
  const bool flip = ((lt0 - m) & 1) != 0;
  __shared__ float Ss[KC][PA], Aa[KC][PA];
  __shared__ float Pe[KC][PB], Po[KC][PB];
  const int wr = wave / 2, wc = wave % 2;
  floatx4 accE[MT][NT], accO[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int c = 0; c < NT; ++c) accE[a][c] = accO[a][c] = floatx4{0, 0, 0, 0};
  const float* pm = ptab + m_off[m];
  const bool odd_eq = (J & 1);
  constexpr int QS = kThreads / KC;
  static_assert(kThreads % KC == 0 && QS % 2 == 0, "row stride must be even");
  const int q0 = tid / KC, kq = tid % KC;
  const float* fbase = nullptr;
  {
    int b, r;
    row_of(m, row0 + q0, &b, &r);
    fbase = fourier + ((int64_t)b * Rrows + r) * J;
  }
  const int64_t fstride = (int64_t)(m == 0 ? QS : QS / 2) * Rrows * J;
  float ru[S_PER], rd[S_PER], rp[P_PER];
  auto fetch = [&](int jj0) {
    const int jj = jj0 + kq;
    const bool eq = odd_eq && jj == 0;
#pragma unroll
    for (int i = 0; i < S_PER; ++i) {
      bool ok = (row0 + q0 + i * QS < nrows) && jj < Jh;
      ru[i] = ok ? fbase[i * fstride + J - Jh + jj] : 0.f;
      rd[i] = (ok && !eq) ? fbase[i * fstride + Jh - 1 - jj] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < P_PER; ++i) {
      int idx = tid + i * kThreads;
      int c = idx / KC, k = idx % KC;
      int l = lt0 + c, jj = jj0 + k;
      rp[i] = (l >= m && l < L && jj < Jhp) ? pm[(int64_t)(l - m) * Jhp + jj] : 0.f;
    }
  };
  fetch(0);
  for (int jj0 = 0; jj0 < Jh; jj0 += KC) {
#pragma unroll
    for (int i = 0; i < S_PER; ++i) {
      int idx = tid + i * kThreads;
      int q = idx / KC, k = idx % KC;
      Ss[k][q] = ru[i] + rd[i];
      Aa[k][q] = ru[i] - rd[i];
    }
#pragma unroll
    for (int i = 0; i < P_PER; ++i) {
      int idx = tid + i * kThreads;
      int c = idx / KC, k = idx % KC;
      if (c & 1) Po[k][c >> 1] = rp[i];
      else Pe[k][c >> 1] = rp[i];
    }
    __syncthreads();
    if (jj0 + KC < Jh) fetch(jj0 + KC);
    const float(*Xe)[PA] = flip ? Aa : Ss;
    const float(*Xo)[PA] = flip ? Ss : Aa;
#pragma unroll
    for (int ks = 0; ks < KC; ks += Pol::KSTEP) {
#pragma unroll
      for (int a = 0; a < MT; ++a) {
#pragma unroll
        for (int c = 0; c < NT; ++c) {
          int ar = wr * WR + a * 16, bc = wc * WC + c * 16;
          accE[a][c] = Pol::template mma<PA, PB>(Xe, ar, Pe, bc, ks, lane, accE[a][c]);
          accO[a][c] = Pol::template mma<PA, PB>(Xo, ar, Po, bc, ks, lane, accO[a][c]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < MT; ++a) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int row = row0 + wr * WR + a * 16 + 4 * (lane / 16) + i;
      if (row >= nrows) continue;
      int b, r;
      row_of(m, row, &b, &r);
      float* out = modal + ((int64_t)b * Rrows + r) * L;
#pragma unroll
      for (int c = 0; c < NT; ++c) {
        int ce = wc * WC + c * 16 + lane % 16;
        int le = lt0 + 2 * ce, lo = le + 1;
        if (le < L) out[le] = accE[a][c][i];
        if (lo < L) out[lo] = accO[a][c][i];
      }
    }
  }
}

}  // namespace

namespace {

template <class Pol, int TC, int KH>
int launch_inv(const dino_plan* pl, const float* modal, float* fourier, int64_t B, hipStream_t s) {
  dim3 grid((pl->Jh + TC - 1) / TC, (unsigned)((2 * B + TR - 1) / TR), pl->M);
  const int mpair = (pl->sched[DINO_STAGE_LEGENDRE_INVERSE] & DINO_SCHED_MPAIRED) ? 1 : 0;
  leg_inv_mfma_kernel<Pol, TC, KH><<<grid, kThreads, 0, s>>>(
      modal, fourier, pl->p32, pl->m_offset, (int)B, pl->L, pl->J, pl->Jh, pl->Jhp, pl->R, mpair);
  DINO_LAUNCH_CHECK("legendre inverse mfma kernel");
  return DINO_OK;
}

template <class Pol, int KC>
int launch_fwd(const dino_plan* pl, const float* fourier, float* modal, int64_t B, hipStream_t s) {
  dim3 grid((pl->L + 127) / 128, (unsigned)((2 * B + TR - 1) / TR), pl->M);
  const int mpair = (pl->sched[DINO_STAGE_LEGENDRE_FORWARD] & DINO_SCHED_MPAIRED) ? 1 : 0;
  leg_fwd_mfma_kernel<Pol, KC><<<grid, kThreads, 0, s>>>(
      fourier, modal, pl->p32, pl->m_offset, (int)B, pl->L, pl->J, pl->Jh, pl->Jhp, pl->R, mpair);
  DINO_LAUNCH_CHECK("legendre forward mfma kernel");
  return DINO_OK;
}

// Node-tile width by latitude count; `narrow` halves the widest tile, which
// doubles the workgroups per (rows, m) at the cost of re-reading the modal
// rows once more.
template <class Pol, int KH>
int inv_by_width(const dino_plan* pl, const float* in, float* out, int64_t B, bool narrow,
                 hipStream_t s) {
  if (pl->Jh > 64 && !narrow) return launch_inv<Pol, (KH > 16 ? 64 : 128), KH>(pl, in, out, B, s);
  if (pl->Jh > 32 && !narrow) return launch_inv<Pol, 64, KH>(pl, in, out, B, s);
  if (pl->Jh > 64) return launch_inv<Pol, (KH > 16 ? 32 : 64), KH>(pl, in, out, B, s);
  return launch_inv<Pol, 32, KH>(pl, in, out, B, s);
}

}  // namespace

int legendre_forward_rows(const dino_plan* pl, const float* in, float* out, int64_t B,
                          bool split, bool triangle_only, bool wide, hipStream_t s);
int legendre_inverse_rows(const dino_plan* pl, const float* in, float* out, int64_t B,
                          bool split, hipStream_t s);

int legendre_inverse_mfma(const dino_plan* pl, const void* modal, void* fourier, int64_t B,
                          int precision, hipStream_t s) {
  if (B == 0) return DINO_OK;
  const float* in = static_cast<const float*>(modal);
  float* out = static_cast<float*>(fourier);
  const int fam = pl->sched[DINO_STAGE_LEGENDRE_INVERSE] & 0xff;
  if (pl->int_nt > 0 && (fam == DINO_SCHED_AUTO || fam == DINO_SCHED_ROWS) &&
      (precision == DINO_PREC_NATIVE || precision == DINO_PREC_BF16X3))
    return legendre_inverse_rows(pl, in, out, B, precision == DINO_PREC_BF16X3, s);
  DINO_REQUIRE((2 * B + TR - 1) / TR <= 65535, DINO_EINVAL, "batch too large");
  const bool narrow = fam == DINO_SCHED_TILED_NARROW;
  switch (precision) {
    case DINO_PREC_NATIVE: return inv_by_width<PolicyF32, 16>(pl, in, out, B, narrow, s);
    case DINO_PREC_BF16X3: return inv_by_width<PolicySplit, 32>(pl, in, out, B, narrow, s);
    default:
      set_error("unknown precision %d", precision);
      return DINO_EINVAL;
  }
}

int legendre_forward_mfma(const dino_plan* pl, const void* fourier, void* modal, int64_t B,
                          int precision, bool triangle_only, hipStream_t s) {
  if (B == 0) return DINO_OK;
  const float* in = static_cast<const float*>(fourier);
  float* out = static_cast<float*>(modal);
  const int fam = pl->sched[DINO_STAGE_LEGENDRE_FORWARD] & 0xff;
  if (pl->fks > 0 &&
      (fam == DINO_SCHED_AUTO || fam == DINO_SCHED_ROWS || fam == DINO_SCHED_ROWS_WIDE) &&
      (precision == DINO_PREC_NATIVE || precision == DINO_PREC_BF16X3))
    return legendre_forward_rows(pl, in, out, B, precision == DINO_PREC_BF16X3, triangle_only,
                                 fam == DINO_SCHED_ROWS_WIDE, s);
  DINO_REQUIRE((2 * B + TR - 1) / TR <= 65535, DINO_EINVAL, "batch too large");
  switch (precision) {
    case DINO_PREC_NATIVE:
      // narrow: 16 nodes per reduction chunk (half the LDS, more resident workgroups)
      if (fam == DINO_SCHED_TILED_NARROW) return launch_fwd<PolicyF32, 16>(pl, in, out, B, s);
      return launch_fwd<PolicyF32, 32>(pl, in, out, B, s);
    case DINO_PREC_BF16X3: return launch_fwd<PolicySplit, 32>(pl, in, out, B, s);
    default:
      set_error("unknown precision %d", precision);
      return DINO_EINVAL;
  }
}

}  // namespace dino
