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
//   split : hi + lo halves of both operands (FP16 or BF16), three
//           v_mfma_f32_16x16x32 products (hi*hi + hi*lo + lo*hi) accumulated
//           in FP32.
#include "dino_common.h"

namespace dino {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
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

template <bool BF>
struct PolicySplit {
  static constexpr int KSTEP = 32;
  template <int PA, int PB>
  __device__ static floatx4 mma(const float (*As)[PA], int arow, const float (*Bs)[PB],
                                int bcol, int k0, int lane, floatx4 acc) {
    const int kb = k0 + 8 * (lane / 16);
    const int ra = arow + lane % 16, cb = bcol + lane % 16;
    if constexpr (BF) {
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
    } else {
      halfx8 ah, al, bh, bl;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        float a = As[kb + i][ra], b = Bs[kb + i][cb];
        _Float16 h = (_Float16)a, g = (_Float16)b;
        ah[i] = h;
        al[i] = (_Float16)(a - (float)h);
        bh[i] = g;
        bl[i] = (_Float16)(b - (float)g);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
      return __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
    }
  }
};

// ------------------------------------------------------------ inverse
// KH degrees per parity per chunk (chunk = 2 KH consecutive degrees).
template <class Pol, int TC, int KH>
__global__ __launch_bounds__(kThreads) void leg_inv_mfma_kernel(
    const float* __restrict__ modal, float* __restrict__ fourier, const float* __restrict__ ptab,
    const int64_t* __restrict__ m_off, const float* __restrict__ scale, int B, int L, int J,
    int Jh, int Jhp, int Rrows) {
  constexpr int WR = TR / 2, WC = TC / 2, MT = WR / 16, NT = WC / 16;
  constexpr int KC = 2 * KH;
  constexpr int PA = TR + 4, PB = TC + 4;
  const int m = blockIdx.z;
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
  for (int l0 = m; l0 < L; l0 += KC) {
    for (int idx = tid; idx < TR * KC; idx += kThreads) {
      int q = idx / KC, k = idx % KC;
      int row = row0 + q, l = l0 + k;
      float v = 0.f;
      if (row < nrows && l < L) {
        int b, r;
        row_of(m, row, &b, &r);
        v = modal[((int64_t)b * Rrows + r) * L + l];
        if (scale) v *= scale[(int64_t)b * Rrows + r];
      }
      if (k & 1) Ao[k >> 1][q] = v;
      else Ae[k >> 1][q] = v;
    }
    for (int idx = tid; idx < KC * TC; idx += kThreads) {
      int k = idx / TC, c = idx % TC;
      int l = l0 + k, jj = jj0 + c;
      float v = (l < L && jj < Jhp) ? pm[(int64_t)(l - m) * Jhp + jj] : 0.f;
      if (k & 1) Po[k >> 1][c] = v;
      else Pe[k >> 1][c] = v;
    }
    __syncthreads();
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
      float inv = scale ? 1.f / scale[(int64_t)b * Rrows + r] : 1.f;
      float* out = fourier + ((int64_t)b * Rrows + r) * J;
#pragma unroll
      for (int c = 0; c < NT; ++c) {
        int jj = jj0 + wc * WC + c * 16 + lane % 16;
        if (jj >= Jh) continue;
        int jup = J - Jh + jj, jdn = Jh - 1 - jj;
        float e = accE[a][c][i], o = accO[a][c][i];
        out[jup] = (e + o) * inv;
        if (jdn != jup) out[jdn] = (e - o) * inv;
      }
    }
  }
}

// ------------------------------------------------------------ forward
// KC latitudes per chunk; tile of 128 degrees starting at l = m + 128 * tile.
template <class Pol, int KC>
__global__ __launch_bounds__(kThreads) void leg_fwd_mfma_kernel(
    const float* __restrict__ fourier, float* __restrict__ modal, const float* __restrict__ ptab,
    const int64_t* __restrict__ m_off, const float* __restrict__ scale, int B, int L, int J,
    int Jh, int Jhp, int Rrows) {
  constexpr int TLH = 64;  // degrees per parity in the tile
  constexpr int WR = TR / 2, WC = TLH / 2, MT = WR / 16, NT = WC / 16;
  constexpr int PA = TR + 4, PB = TLH + 4;
  const int m = blockIdx.z;
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.y * TR;
  if (row0 >= nrows) return;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int lt0 = m + blockIdx.x * (2 * TLH);
  if (blockIdx.x == 0 && m > 0) {
    // structural zeros for l < m
    for (int idx = tid; idx < TR * m; idx += kThreads) {
      int q = idx / m, l = idx % m;
      int row = row0 + q;
      if (row < nrows) {
        int b, r;
        row_of(m, row, &b, &r);
        modal[((int64_t)b * Rrows + r) * L + l] = 0.f;
      }
    }
  }
  if (lt0 >= L) return;
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
  for (int jj0 = 0; jj0 < Jh; jj0 += KC) {
    for (int idx = tid; idx < TR * KC; idx += kThreads) {
      int q = idx / KC, k = idx % KC;
      int row = row0 + q, jj = jj0 + k;
      float sv = 0.f, av = 0.f;
      if (row < nrows && jj < Jh) {
        int b, r;
        row_of(m, row, &b, &r);
        const float* f = fourier + ((int64_t)b * Rrows + r) * J;
        int jup = J - Jh + jj, jdn = Jh - 1 - jj;
        float u = f[jup];
        if (odd_eq && jj == 0) {
          sv = u;
        } else {
          float d = f[jdn];
          sv = u + d;
          av = u - d;
        }
        if (scale) {
          float s = scale[(int64_t)b * Rrows + r];
          sv *= s;
          av *= s;
        }
      }
      Ss[k][q] = sv;
      Aa[k][q] = av;
    }
    for (int idx = tid; idx < 2 * TLH * KC; idx += kThreads) {
      int c = idx / KC, k = idx % KC;
      int l = lt0 + c, jj = jj0 + k;
      float v = (l < L && jj < Jhp) ? pm[(int64_t)(l - m) * Jhp + jj] : 0.f;
      if (c & 1) Po[k][c >> 1] = v;
      else Pe[k][c >> 1] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC; ks += Pol::KSTEP) {
#pragma unroll
      for (int a = 0; a < MT; ++a) {
#pragma unroll
        for (int c = 0; c < NT; ++c) {
          int ar = wr * WR + a * 16, bc = wc * WC + c * 16;
          accE[a][c] = Pol::template mma<PA, PB>(Ss, ar, Pe, bc, ks, lane, accE[a][c]);
          accO[a][c] = Pol::template mma<PA, PB>(Aa, ar, Po, bc, ks, lane, accO[a][c]);
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
      float inv = scale ? 1.f / scale[(int64_t)b * Rrows + r] : 1.f;
      float* out = modal + ((int64_t)b * Rrows + r) * L;
#pragma unroll
      for (int c = 0; c < NT; ++c) {
        int ce = wc * WC + c * 16 + lane % 16;
        int le = lt0 + 2 * ce, lo = le + 1;
        if (le < L) out[le] = accE[a][c][i] * inv;
        if (lo < L) out[lo] = accO[a][c][i] * inv;
      }
    }
  }
}

template <class Pol, int TC, int KH>
int launch_inv(const dino_plan* pl, const float* modal, float* fourier, int64_t B, hipStream_t s) {
  dim3 grid((pl->Jh + TC - 1) / TC, (unsigned)((2 * B + TR - 1) / TR), pl->M);
  leg_inv_mfma_kernel<Pol, TC, KH><<<grid, kThreads, 0, s>>>(
      modal, fourier, pl->p32, pl->m_offset, nullptr, (int)B, pl->L, pl->J, pl->Jh, pl->Jhp,
      pl->R);
  DINO_LAUNCH_CHECK("legendre inverse mfma kernel");
  return DINO_OK;
}

template <class Pol, int KC>
int launch_fwd(const dino_plan* pl, const float* fourier, float* modal, int64_t B, hipStream_t s) {
  dim3 grid((pl->L + 127) / 128, (unsigned)((2 * B + TR - 1) / TR), pl->M);
  leg_fwd_mfma_kernel<Pol, KC><<<grid, kThreads, 0, s>>>(
      fourier, modal, pl->p32, pl->m_offset, nullptr, (int)B, pl->L, pl->J, pl->Jh, pl->Jhp,
      pl->R);
  DINO_LAUNCH_CHECK("legendre forward mfma kernel");
  return DINO_OK;
}

template <class Pol, int KH>
int inv_by_width(const dino_plan* pl, const float* in, float* out, int64_t B, hipStream_t s) {
  if (pl->Jh > 64) return launch_inv<Pol, (KH > 8 ? 64 : 128), KH>(pl, in, out, B, s);
  if (pl->Jh > 32) return launch_inv<Pol, 64, KH>(pl, in, out, B, s);
  return launch_inv<Pol, 32, KH>(pl, in, out, B, s);
}

}  // namespace

int legendre_inverse_mfma(const dino_plan* pl, const void* modal, void* fourier, int64_t B,
                          int precision, hipStream_t s) {
  if (B == 0) return DINO_OK;
  DINO_REQUIRE((2 * B + TR - 1) / TR <= 65535, DINO_EINVAL, "batch too large");
  const float* in = static_cast<const float*>(modal);
  float* out = static_cast<float*>(fourier);
  switch (precision) {
    case DINO_PREC_NATIVE: return inv_by_width<PolicyF32, 8>(pl, in, out, B, s);
    case DINO_PREC_BF16X3: return inv_by_width<PolicySplit<true>, 32>(pl, in, out, B, s);
    default:
      set_error("split precision %d is not available; use DINO_PREC_BF16X3", precision);
      return DINO_EINVAL;
  }
}

int legendre_forward_mfma(const dino_plan* pl, const void* fourier, void* modal, int64_t B,
                          int precision, hipStream_t s) {
  if (B == 0) return DINO_OK;
  DINO_REQUIRE((2 * B + TR - 1) / TR <= 65535, DINO_EINVAL, "batch too large");
  const float* in = static_cast<const float*>(fourier);
  float* out = static_cast<float*>(modal);
  switch (precision) {
    case DINO_PREC_NATIVE: return launch_fwd<PolicyF32, 16>(pl, in, out, B, s);
    case DINO_PREC_BF16X3: return launch_fwd<PolicySplit<true>, 32>(pl, in, out, B, s);
    default:
      set_error("split precision %d is not available; use DINO_PREC_BF16X3", precision);
      return DINO_EINVAL;
  }
}

}  // namespace dino
