// Legendre transform: per zonal wavenumber m, a GEMM on the triangular slab
// l >= m against the folded P table, with the equatorial symmetry
// P_l^m(-x) = (-1)^(l+m) P_l^m(x) applied in the tile loads / epilogue.
//
// inverse: fourier[b, r, j] = sum_l P[m, j, l] * modal[b, r, l]
//          rows (b, r) with r in {2m-1, 2m} share P[m]; even (l+m) and odd
//          (l+m) partial sums E, O give j_up = E + O and j_dn = E - O.
// forward: modal[b, r, l] = sum_j P[m, j, l] * fourier[b, r, j]
//          folded to S = f[j_up] + f[j_dn] (even l+m) or A = f[j_up] - f[j_dn]
//          (odd l+m) over the Jh northern nodes.
// These replace the dense einsums 'mjl,...ml->...mj' and 'mjl,...mj->...ml'
// (spherical_harmonic.py:265, 277-279), which multiply by the l < m zeros and
// the duplicated +/-m rows of the reference's (2M-1, J, L) table.
//
// This file holds the plain FMA kernels (FP32 or FP64 arithmetic); the split
// FP16/BF16 MFMA kernels live in legendre_mfma.hip.
#include "dino_common.h"

namespace dino {
namespace {

constexpr int TR = 64;   // rows (fields x cos/sin) per tile
constexpr int TC = 64;   // latitudes (inverse) or degrees (forward) per tile
constexpr int KC = 16;   // contraction chunk
constexpr int NT = 256;  // threads: 16 x 16, each a 4 x 4 micro-tile

__device__ __forceinline__ void row_of(int m, int q, int* b, int* r) {
  if (m == 0) {
    *b = q;
    *r = 0;
  } else {
    *b = q >> 1;
    *r = 2 * m - 1 + (q & 1);
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void legendre_inverse_fma_kernel(
    const T* __restrict__ modal, T* __restrict__ fourier, const T* __restrict__ ptab,
    const int64_t* __restrict__ m_off, int B, int L, int J, int Jh, int Jhp, int Rrows) {
  const int m = blockIdx.z;
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.y * TR;
  if (row0 >= nrows) return;
  const int jj0 = blockIdx.x * TC;
  __shared__ T As[KC][TR + 4];
  __shared__ T Ps[KC][TC + 4];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  T accE[4][4], accO[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) accE[i][c] = accO[i][c] = 0;
  const T* pm = ptab + m_off[m];
  for (int l0 = m; l0 < L; l0 += KC) {
    // A tile: rows x KC degrees (modal rows are contiguous in l)
    for (int idx = tid; idx < TR * KC; idx += NT) {
      int q = idx / KC, k = idx % KC;
      int row = row0 + q, l = l0 + k;
      T v = 0;
      if (row < nrows && l < L) {
        int b, r;
        row_of(m, row, &b, &r);
        v = modal[((int64_t)b * Rrows + r) * L + l];
      }
      As[k][q] = v;
    }
    // P tile: KC degrees x TC latitudes
    for (int idx = tid; idx < KC * TC; idx += NT) {
      int k = idx / TC, c = idx % TC;
      int l = l0 + k, jj = jj0 + c;
      Ps[k][c] = (l < L && jj < Jhp) ? pm[(int64_t)(l - m) * Jhp + jj] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KC; k += 2) {
      T a0[4], a1[4], p0[4], p1[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a0[i] = As[k][ty * 4 + i];
        a1[i] = As[k + 1][ty * 4 + i];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        p0[c] = Ps[k][tx * 4 + c];
        p1[c] = Ps[k + 1][tx * 4 + c];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          accE[i][c] += a0[i] * p0[c];
          accO[i][c] += a1[i] * p1[c];
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int row = row0 + ty * 4 + i;
    if (row >= nrows) continue;
    int b, r;
    row_of(m, row, &b, &r);
    T* out = fourier + ((int64_t)b * Rrows + r) * J;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int jj = jj0 + tx * 4 + c;
      if (jj >= Jh) continue;
      int jup = J - Jh + jj;
      int jdn = Jh - 1 - jj;
      out[jup] = accE[i][c] + accO[i][c];
      if (jdn != jup) out[jdn] = accE[i][c] - accO[i][c];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void legendre_forward_fma_kernel(
    const T* __restrict__ fourier, T* __restrict__ modal, const T* __restrict__ ptab,
    const int64_t* __restrict__ m_off, int B, int L, int J, int Jh, int Jhp, int Rrows) {
  const int m = blockIdx.z;
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.y * TR;
  if (row0 >= nrows) return;
  const int lt0 = blockIdx.x * TC;  // first degree of this tile
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  if (lt0 + TC <= m) {
    // whole tile below the diagonal: structural zeros
    for (int idx = tid; idx < TR * TC; idx += NT) {
      int q = idx / TC, c = idx % TC;
      int row = row0 + q, l = lt0 + c;
      if (row < nrows && l < L) {
        int b, r;
        row_of(m, row, &b, &r);
        modal[((int64_t)b * Rrows + r) * L + l] = 0;
      }
    }
    return;
  }
  __shared__ T Ss[KC][TR + 4];
  __shared__ T As[KC][TR + 4];
  __shared__ T Ps[KC][TC + 4];
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = 0;
  const T* pm = ptab + m_off[m];
  const bool odd_eq = (J & 1);
  for (int jj0 = 0; jj0 < Jh; jj0 += KC) {
    for (int idx = tid; idx < TR * KC; idx += NT) {
      int q = idx / KC, k = idx % KC;
      int row = row0 + q, jj = jj0 + k;
      T s = 0, a = 0;
      if (row < nrows && jj < Jh) {
        int b, r;
        row_of(m, row, &b, &r);
        const T* f = fourier + ((int64_t)b * Rrows + r) * J;
        int jup = J - Jh + jj, jdn = Jh - 1 - jj;
        T u = f[jup];
        if (odd_eq && jj == 0) {
          s = u;
          a = 0;
        } else {
          T d = f[jdn];
          s = u + d;
          a = u - d;
        }
      }
      Ss[k][q] = s;
      As[k][q] = a;
    }
    for (int idx = tid; idx < KC * TC; idx += NT) {
      int k = idx / TC, c = idx % TC;
      int jj = jj0 + k, l = lt0 + c;
      Ps[k][c] = (jj < Jhp && l >= m && l < L) ? pm[(int64_t)(l - m) * Jhp + jj] : (T)0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < KC; ++k) {
      T s[4], a[4], p[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        s[i] = Ss[k][ty * 4 + i];
        a[i] = As[k][ty * 4 + i];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) p[c] = Ps[k][tx * 4 + c];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        // degree l = lt0 + tx*4 + c; parity of (l - m) selects S or A
        bool even = (((lt0 + tx * 4 + c - m) & 1) == 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][c] += (even ? s[i] : a[i]) * p[c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int row = row0 + ty * 4 + i;
    if (row >= nrows) continue;
    int b, r;
    row_of(m, row, &b, &r);
    T* out = modal + ((int64_t)b * Rrows + r) * L;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int l = lt0 + tx * 4 + c;
      if (l < L) out[l] = (l >= m) ? acc[i][c] : (T)0;
    }
  }
}

template <typename T>
int run_legendre(bool forward, const dino_plan* pl, const void* in, void* out, int64_t B,
                 hipStream_t s) {
  if (B == 0) return DINO_OK;
  DINO_REQUIRE(2 * B <= (int64_t)65535 * TR, DINO_EINVAL, "batch too large");
  const T* ptab;
  if constexpr (sizeof(T) == 8) ptab = pl->p64;
  else ptab = pl->p32;
  int row_tiles = (int)((2 * B + TR - 1) / TR);
  if (forward) {
    dim3 grid((pl->L + TC - 1) / TC, row_tiles, pl->M);
    legendre_forward_fma_kernel<T><<<grid, NT, 0, s>>>(
        static_cast<const T*>(in), static_cast<T*>(out), ptab, pl->m_offset, (int)B, pl->L,
        pl->J, pl->Jh, pl->Jhp, pl->R);
  } else {
    dim3 grid((pl->Jh + TC - 1) / TC, row_tiles, pl->M);
    legendre_inverse_fma_kernel<T><<<grid, NT, 0, s>>>(
        static_cast<const T*>(in), static_cast<T*>(out), ptab, pl->m_offset, (int)B, pl->L,
        pl->J, pl->Jh, pl->Jhp, pl->R);
  }
  DINO_LAUNCH_CHECK("legendre fma kernel");
  return DINO_OK;
}

}  // namespace

int legendre_inverse_fma(const dino_plan* pl, const void* modal, void* fourier, int64_t B,
                         int dtype, hipStream_t s) {
  if (dtype == DINO_F64) return run_legendre<double>(false, pl, modal, fourier, B, s);
  return run_legendre<float>(false, pl, modal, fourier, B, s);
}

int legendre_forward_fma(const dino_plan* pl, const void* fourier, void* modal, int64_t B,
                         int dtype, hipStream_t s) {
  if (dtype == DINO_F64) return run_legendre<double>(true, pl, fourier, modal, B, s);
  return run_legendre<float>(true, pl, fourier, modal, B, s);
}

}  // namespace dino
