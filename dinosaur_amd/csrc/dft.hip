// Longitude transform: batched real DFT over (fields x latitudes).
//
// Power-of-two I: a workgroup owns one field b and a tile of 2P consecutive
// latitude columns.  Column pairs (a, b) are packed as z = x_a + i x_b and one
// complex FFT of length I per pair runs in LDS as Stockham radix-8/4/2
// passes, with the twiddles exp(-2 pi i k / I) staged in LDS.  HBM traffic is
// one coalesced read of the nodal tile and one write of the 2M-1 Fourier rows
// (the algorithmic minimum), versus the reference's dense (I x 2M-1) matmul
// (spherical_harmonic.py:266-268, 275-276).
//
// Other I: a direct real DFT with the cos/sin table in LDS (O(I M) per column).
//
// Real-basis conventions (fourier.py:24-60): row 0 is 1/sqrt(2 pi), row 2m-1
// is cos(m lambda)/sqrt(pi), row 2m is sin(m lambda)/sqrt(pi).
#include <algorithm>
#include <type_traits>

#include "dino_common.h"

namespace dino {
namespace {

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) {
  return INV ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
  cplx<T> s02 = cadd(a0, a2), d02 = csub(a0, a2);
  cplx<T> s13 = cadd(a1, a3), d13 = mul_mi<T, INV>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4v(cplx<T>* v) { dft4<T, INV>(v[0], v[1], v[2], v[3]); }

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  const T c = (T)0.70710678118654752440;
  cplx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cplx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<T, INV>(e0, e1, e2, e3);
  dft4<T, INV>(o0, o1, o2, o3);
  // W8^k for k = 1, 2, 3 (conjugated for the inverse)
  cplx<T> w1 = INV ? cplx<T>{c, c} : cplx<T>{c, -c};
  cplx<T> w3 = INV ? cplx<T>{-c, c} : cplx<T>{-c, -c};
  o1 = cmul(o1, w1);
  o2 = mul_mi<T, INV>(o2);
  o3 = cmul(o3, w3);
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_r(cplx<T>* v) {
  if constexpr (R == 2) dft2<T, INV>(v);
  else if constexpr (R == 4) dft4v<T, INV>(v);
  else dft8<T, INV>(v);
}

// One in-place Stockham pass on P interleaved complex columns of length N:
// buf[n * P + p].  All butterflies of the pass are loaded to registers before
// any store (barrier in between), so one LDS buffer suffices.
template <typename T, int N, int P, int NT, int R, bool INV>
__device__ __forceinline__ void stockham_pass(cplx<T>* buf, const cplx<T>* tws, int Ns) {
  constexpr int NB = N / R;
  constexpr int TOT = NB * P;
  constexpr int BPT = (TOT + NT - 1) / NT;
  cplx<T> v[BPT][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    int t = tid + q * NT;
    if (TOT % NT == 0 || t < TOT) {
      int p = t % P, j = t / P;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[(j + r * NB) * P + p];
      if (Ns > 1) {
        int k = j % Ns;
        int step = N / (Ns * R);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> w = tws[r * k * step];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_r<T, R, INV>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    int t = tid + q * NT;
    if (TOT % NT == 0 || t < TOT) {
      int p = t % P, j = t / P;
      int k = j % Ns;
      int d = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(d + r * Ns) * P + p] = v[q][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int P, int NT, bool INV>
__device__ __forceinline__ void fft_lds(cplx<T>* buf, const cplx<T>* tws) {
  int Ns = 1;
  int rem = N;
  while (rem >= 8) {
    stockham_pass<T, N, P, NT, 8, INV>(buf, tws, Ns);
    Ns *= 8;
    rem /= 8;
  }
  if (rem == 4) {
    stockham_pass<T, N, P, NT, 4, INV>(buf, tws, Ns);
  } else if (rem == 2) {
    stockham_pass<T, N, P, NT, 2, INV>(buf, tws, Ns);
  }
}

// Tile (N rows x C columns starting at column j0) <-> LDS [n][c].  When the
// tile is fully inside J and 16-byte aligned, every thread issues all of its
// vector loads before any LDS store (compile-time trip count), so a
// workgroup keeps its whole tile in flight; otherwise a guarded scalar loop.
template <typename T, int N, int C, int NT>
__device__ __forceinline__ void load_tile(const T* __restrict__ src, int J, int j0,
                                          const T* __restrict__ qw, T* lds) {
  constexpr int V = 16 / sizeof(T);  // elements per 16-byte vector
  const int tid = threadIdx.x;
  if constexpr (C % V == 0 && (N * C / V) % NT == 0) {
    if (j0 + C <= J && (J % V) == 0) {
      using VT = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
      constexpr int CV = C / V;
      constexpr int IT = N * CV / NT;
      static_assert(NT % CV == 0, "column vector index must be fixed per thread");
      VT v[IT];
      const int cv = tid % CV;  // fixed across iterations since NT % CV == 0
#pragma unroll
      for (int q = 0; q < IT; ++q) {
        int idx = tid + q * NT;
        int n = idx / CV;
        v[q] = *reinterpret_cast<const VT*>(src + (int64_t)n * J + j0 + cv * V);
      }
      T w[V];
#pragma unroll
      for (int e = 0; e < V; ++e) w[e] = qw ? qw[j0 + cv * V + e] : (T)1;
#pragma unroll
      for (int q = 0; q < IT; ++q) {
        int idx = tid + q * NT;
        int n = idx / CV;
        const T* pv = reinterpret_cast<const T*>(&v[q]);
#pragma unroll
        for (int e = 0; e < V; ++e) lds[n * C + cv * V + e] = pv[e] * w[e];
      }
      return;
    }
  }
  for (int idx = tid; idx < N * C; idx += NT) {
    int n = idx / C, c = idx % C;
    int j = j0 + c;
    T v = 0;
    if (j < J) {
      v = src[(int64_t)n * J + j];
      if (qw) v *= qw[j];
    }
    lds[n * C + c] = v;
  }
}

template <typename T, int N, int C, int NT>
__device__ __forceinline__ void store_tile(const T* lds, T* __restrict__ dst, int J, int j0,
                                           const T* __restrict__ qw) {
  constexpr int V = 16 / sizeof(T);
  const int tid = threadIdx.x;
  if constexpr (C % V == 0 && (N * C / V) % NT == 0) {
    if (j0 + C <= J && (J % V) == 0) {
      using VT = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
      constexpr int CV = C / V;
      constexpr int IT = N * CV / NT;
      static_assert(NT % CV == 0, "column vector index must be fixed per thread");
      const int cv = tid % CV;
      T w[V];
#pragma unroll
      for (int e = 0; e < V; ++e) w[e] = qw ? qw[j0 + cv * V + e] : (T)1;
#pragma unroll
      for (int q = 0; q < IT; ++q) {
        int idx = tid + q * NT;
        int n = idx / CV;
        VT v;
        T* pv = reinterpret_cast<T*>(&v);
#pragma unroll
        for (int e = 0; e < V; ++e) pv[e] = lds[n * C + cv * V + e] * w[e];
        *reinterpret_cast<VT*>(dst + (int64_t)n * J + j0 + cv * V) = v;
      }
      return;
    }
  }
  for (int idx = tid; idx < N * C; idx += NT) {
    int n = idx / C, c = idx % C;
    int j = j0 + c;
    if (j >= J) continue;
    T v = lds[n * C + c];
    if (qw) v *= qw[j];
    dst[(int64_t)n * J + j] = v;
  }
}

// nodal (B, N, J) -> fourier (B, R, J)
template <typename T, int N, int P, int NT>
__global__ __launch_bounds__(NT) void dft_forward_fft_kernel(
    const T* __restrict__ nodal, T* __restrict__ fourier, const T* __restrict__ tw,
    const T* __restrict__ qw, int J, int Rrows, int weighted) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tws = buf + N * P;
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.y;
  const int j0 = blockIdx.x * (2 * P);
  for (int k = tid; k < N; k += NT) tws[k] = cplx<T>{tw[2 * k], -tw[2 * k + 1]};
  const T* src = nodal + b * (int64_t)N * J;
  T* bufr = reinterpret_cast<T*>(buf);
  // (re, im) of pair c/2 interleave as x_a, x_b
  load_tile<T, N, 2 * P, NT>(src, J, j0, weighted ? qw : nullptr, bufr);
  __syncthreads();
  fft_lds<T, N, P, NT, false>(buf, tws);
  const T inv_sqrt_2pi = (T)0.39894228040143267794;
  const T inv_sqrt_pi = (T)0.56418958354775628695;
  T* dst = fourier + b * (int64_t)Rrows * J;
  for (int idx = tid; idx < Rrows * 2 * P; idx += NT) {
    int r = idx / (2 * P), c = idx % (2 * P);
    int j = j0 + c;
    if (j >= J) continue;
    int m = (r + 1) / 2;
    cplx<T> zm = buf[m * P + c / 2];
    cplx<T> zn = buf[((N - m) % N) * P + c / 2];
    T re, im;
    if ((c & 1) == 0) {
      re = (zm.x + zn.x) * (T)0.5;
      im = (zm.y - zn.y) * (T)0.5;
    } else {
      re = (zm.y + zn.y) * (T)0.5;
      im = -(zm.x - zn.x) * (T)0.5;
    }
    T out;
    if (r == 0) out = re * inv_sqrt_2pi;
    else if (r & 1) out = re * inv_sqrt_pi;
    else out = -im * inv_sqrt_pi;
    dst[(int64_t)r * J + j] = out;
  }
}

// fourier (B, R, J) -> nodal (B, N, J)
template <typename T, int N, int P, int NT>
__global__ __launch_bounds__(NT) void dft_inverse_fft_kernel(
    const T* __restrict__ fourier, T* __restrict__ nodal, const T* __restrict__ tw,
    const T* __restrict__ qw, int J, int M, int weighted) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tws = buf + N * P;
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.y;
  const int j0 = blockIdx.x * (2 * P);
  const int Rrows = 2 * M - 1;
  for (int k = tid; k < N; k += NT) tws[k] = cplx<T>{tw[2 * k], -tw[2 * k + 1]};
  for (int idx = tid; idx < N * P; idx += NT) buf[idx] = cplx<T>{0, 0};
  __syncthreads();
  const T* src = fourier + b * (int64_t)Rrows * J;
  const T inv_sqrt_2pi = (T)0.39894228040143267794;
  const T half_inv_sqrt_pi = (T)0.28209479177387814347;
  for (int idx = tid; idx < M * P; idx += NT) {
    int m = idx / P, p = idx % P;
    int ja = j0 + 2 * p, jb = ja + 1;
    if (m == 0) {
      T ga = ja < J ? src[ja] : (T)0;
      T gb = jb < J ? src[jb] : (T)0;
      buf[p] = cplx<T>{ga * inv_sqrt_2pi, gb * inv_sqrt_2pi};
    } else {
      const T* r1 = src + (int64_t)(2 * m - 1) * J;
      const T* r2 = src + (int64_t)(2 * m) * J;
      T ga1 = ja < J ? r1[ja] : (T)0, ga2 = ja < J ? r2[ja] : (T)0;
      T gb1 = jb < J ? r1[jb] : (T)0, gb2 = jb < J ? r2[jb] : (T)0;
      buf[m * P + p] = cplx<T>{(ga1 + gb2) * half_inv_sqrt_pi, (gb1 - ga2) * half_inv_sqrt_pi};
      buf[(N - m) * P + p] = cplx<T>{(ga1 - gb2) * half_inv_sqrt_pi, (ga2 + gb1) * half_inv_sqrt_pi};
    }
  }
  __syncthreads();
  fft_lds<T, N, P, NT, true>(buf, tws);
  T* dst = nodal + b * (int64_t)N * J;
  store_tile<T, N, 2 * P, NT>(reinterpret_cast<const T*>(buf), dst, J, j0,
                              weighted ? qw : nullptr);
}

// ---------------------------------------------------------------- persistent FP32 FFT
// The workhorse shape (FP32, power-of-two I <= 512, J % 32 == 0, 16-byte
// aligned buffers): persistent workgroups walk the (field, 32-column) tiles
// and prefetch the next tile's input into registers while the current tile
// is transformed in LDS, so HBM reads overlap the FFT passes and the stores
// instead of alternating with them.  Same packing, passes and unpacking as the
// kernels above, hence the same results.
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kFftThreads = 256;
constexpr int kPersistThreads = 512;  // halves the per-thread prefetch and pass registers

template <int N>
__global__ __launch_bounds__(kPersistThreads) void fft_forward_persistent_kernel(
    const float* __restrict__ nodal, float* __restrict__ fourier, const float* __restrict__ tw,
    const float* __restrict__ qw, int J, int Rrows, int weighted, int64_t ntiles, int tpf) {
  constexpr int P = 16, NT = kPersistThreads, CV = 8, ROWS = NT / CV, LPT = N / ROWS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<float>* buf = reinterpret_cast<cplx<float>*>(smem_raw);
  cplx<float>* tws = buf + N * P;
  const int tid = threadIdx.x, cv = tid % CV, n0 = tid / CV;
  for (int k = tid; k < N; k += NT) tws[k] = cplx<float>{tw[2 * k], -tw[2 * k + 1]};
  int64_t tile = blockIdx.x;
  if (tile >= ntiles) return;
  f4v pre[LPT];
  auto issue = [&](int64_t t) {
    const int64_t b = t / tpf;
    const int j0 = (int)(t - b * tpf) * 32;
    const float* src = nodal + b * (int64_t)N * J + j0 + 4 * cv;
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      pre[q] = *reinterpret_cast<const f4v*>(src + (int64_t)(n0 + ROWS * q) * J);
  };
  issue(tile);
  const float inv_sqrt_2pi = 0.39894228040143267794f;
  const float inv_sqrt_pi = 0.56418958354775628695f;
  for (; tile < ntiles; tile += gridDim.x) {
    const int64_t b = tile / tpf;
    const int j0 = (int)(tile - b * tpf) * 32;
    f4v w = {1.f, 1.f, 1.f, 1.f};
    if (weighted) w = *reinterpret_cast<const f4v*>(qw + j0 + 4 * cv);
    // column pair (2p, 2p + 1) -> complex (x_a, x_b): a float4 is two pairs
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      *reinterpret_cast<f4v*>(&buf[(n0 + ROWS * q) * P + 2 * cv]) = pre[q] * w;
    __syncthreads();
    if (tile + gridDim.x < ntiles) issue(tile + gridDim.x);
    fft_lds<float, N, P, NT, false>(buf, tws);
    float* dst = fourier + b * (int64_t)Rrows * J + j0;
    for (int idx = tid; idx < Rrows * CV; idx += NT) {
      const int r = idx / CV, c4 = idx % CV;
      const int m = (r + 1) / 2;
      f4v o;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = 4 * c4 + e;
        const cplx<float> zm = buf[m * P + c / 2];
        const cplx<float> zn = buf[((N - m) % N) * P + c / 2];
        float re, im;
        if ((c & 1) == 0) {
          re = (zm.x + zn.x) * 0.5f;
          im = (zm.y - zn.y) * 0.5f;
        } else {
          re = (zm.y + zn.y) * 0.5f;
          im = -(zm.x - zn.x) * 0.5f;
        }
        o[e] = r == 0 ? re * inv_sqrt_2pi : ((r & 1) ? re * inv_sqrt_pi : -im * inv_sqrt_pi);
      }
      *reinterpret_cast<f4v*>(dst + (int64_t)r * J + 4 * c4) = o;
    }
    __syncthreads();  // buf is rewritten by the next tile
  }
}

template <int N>
__global__ __launch_bounds__(kPersistThreads) void fft_inverse_persistent_kernel(
    const float* __restrict__ fourier, float* __restrict__ nodal, const float* __restrict__ tw,
    const float* __restrict__ qw, int J, int M, int weighted, int64_t ntiles, int tpf) {
  constexpr int P = 16, NT = kPersistThreads, CV = 8, ROWS = NT / CV, LPT = N / ROWS;
  constexpr int MAXM = N / 2;                     // M <= N / 2 + 1 rows of input are packed
  constexpr int IPT = (MAXM * CV + NT - 1) / NT;  // (m, column quad) items per thread
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<float>* buf = reinterpret_cast<cplx<float>*>(smem_raw);
  cplx<float>* tws = buf + N * P;
  const int tid = threadIdx.x;
  const int Rrows = 2 * M - 1;
  for (int k = tid; k < N; k += NT) tws[k] = cplx<float>{tw[2 * k], -tw[2 * k + 1]};
  int64_t tile = blockIdx.x;
  if (tile >= ntiles) return;
  f4v g1[IPT], g2[IPT];
  auto issue = [&](int64_t t) {
    const int64_t b = t / tpf;
    const int j0 = (int)(t - b * tpf) * 32;
    const float* src = fourier + b * (int64_t)Rrows * J + j0;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int idx = tid + NT * i;
      const int m = idx / CV, c4 = idx % CV;
      const bool ok = m < M;
      const int r1 = ok ? (m == 0 ? 0 : 2 * m - 1) : 0;
      const int r2 = ok ? (m == 0 ? 0 : 2 * m) : 0;
      g1[i] = *reinterpret_cast<const f4v*>(src + (int64_t)r1 * J + 4 * c4);
      g2[i] = *reinterpret_cast<const f4v*>(src + (int64_t)r2 * J + 4 * c4);
    }
  };
  issue(tile);
  const float inv_sqrt_2pi = 0.39894228040143267794f;
  const float half_inv_sqrt_pi = 0.28209479177387814347f;
  for (; tile < ntiles; tile += gridDim.x) {
    const int64_t b = tile / tpf;
    const int j0 = (int)(tile - b * tpf) * 32;
    // rows M .. N - M of the spectrum are zero
    for (int idx = tid; idx < (N - 2 * M + 1) * P; idx += NT)
      buf[(M + idx / P) * P + idx % P] = cplx<float>{0.f, 0.f};
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int idx = tid + NT * i;
      const int m = idx / CV, c4 = idx % CV;
      if (m >= M) continue;
      const f4v a = g1[i], c = g2[i];
      if (m == 0) {
        *reinterpret_cast<f4v*>(&buf[2 * c4]) = a * inv_sqrt_2pi;
      } else {
        // pairs p = 2 c4 (a.x, a.y | c.x, c.y) and 2 c4 + 1 (a.z, a.w | c.z, c.w)
        const f4v lo = {(a[0] + c[1]) * half_inv_sqrt_pi, (a[1] - c[0]) * half_inv_sqrt_pi,
                        (a[2] + c[3]) * half_inv_sqrt_pi, (a[3] - c[2]) * half_inv_sqrt_pi};
        const f4v hi = {(a[0] - c[1]) * half_inv_sqrt_pi, (c[0] + a[1]) * half_inv_sqrt_pi,
                        (a[2] - c[3]) * half_inv_sqrt_pi, (c[2] + a[3]) * half_inv_sqrt_pi};
        *reinterpret_cast<f4v*>(&buf[m * P + 2 * c4]) = lo;
        *reinterpret_cast<f4v*>(&buf[(N - m) * P + 2 * c4]) = hi;
      }
    }
    __syncthreads();
    if (tile + gridDim.x < ntiles) issue(tile + gridDim.x);
    fft_lds<float, N, P, NT, true>(buf, tws);
    float* dst = nodal + b * (int64_t)N * J + j0;
    const int cv = tid % CV, n0 = tid / CV;
    f4v w = {1.f, 1.f, 1.f, 1.f};
    if (weighted) w = *reinterpret_cast<const f4v*>(qw + j0 + 4 * cv);
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int n = n0 + ROWS * q;
      *reinterpret_cast<f4v*>(dst + (int64_t)n * J + 4 * cv) =
          *reinterpret_cast<const f4v*>(&buf[n * P + 2 * cv]) * w;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- direct DFT
template <typename T>
__global__ void dft_forward_direct_kernel(const T* __restrict__ nodal, T* __restrict__ fourier,
                                          const T* __restrict__ tw, const T* __restrict__ qw,
                                          int N, int J, int Rrows, int TJ, int weighted) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* cs = reinterpret_cast<T*>(smem_raw);  // (N, 2)
  T* x = cs + 2 * N;                         // (N, TJ)
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t b = blockIdx.y;
  const int j0 = blockIdx.x * TJ;
  for (int k = tid; k < 2 * N; k += nt) cs[k] = tw[k];
  const T* src = nodal + b * (int64_t)N * J;
  for (int idx = tid; idx < N * TJ; idx += nt) {
    int n = idx / TJ, c = idx % TJ, j = j0 + c;
    T v = 0;
    if (j < J) {
      v = src[(int64_t)n * J + j];
      if (weighted) v *= qw[j];
    }
    x[idx] = v;
  }
  __syncthreads();
  const T inv_sqrt_2pi = (T)0.39894228040143267794;
  const T inv_sqrt_pi = (T)0.56418958354775628695;
  T* dst = fourier + b * (int64_t)Rrows * J;
  for (int idx = tid; idx < Rrows * TJ; idx += nt) {
    int r = idx / TJ, c = idx % TJ, j = j0 + c;
    if (j >= J) continue;
    int m = (r + 1) / 2;
    int use_sin = (r > 0) && !(r & 1);
    T acc = 0;
    int k = 0;
    for (int n = 0; n < N; ++n) {
      acc += x[n * TJ + c] * cs[2 * k + use_sin];
      k += m;
      if (k >= N) k -= N;
    }
    dst[(int64_t)r * J + j] = acc * (r == 0 ? inv_sqrt_2pi : inv_sqrt_pi);
  }
}

template <typename T>
__global__ void dft_inverse_direct_kernel(const T* __restrict__ fourier, T* __restrict__ nodal,
                                          const T* __restrict__ tw, const T* __restrict__ qw,
                                          int N, int J, int Rrows, int TJ, int weighted) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* cs = reinterpret_cast<T*>(smem_raw);  // (N, 2)
  T* g = cs + 2 * N;                         // (Rrows, TJ)
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t b = blockIdx.y;
  const int j0 = blockIdx.x * TJ;
  for (int k = tid; k < 2 * N; k += nt) cs[k] = tw[k];
  const T* src = fourier + b * (int64_t)Rrows * J;
  for (int idx = tid; idx < Rrows * TJ; idx += nt) {
    int r = idx / TJ, c = idx % TJ, j = j0 + c;
    g[idx] = j < J ? src[(int64_t)r * J + j] : (T)0;
  }
  __syncthreads();
  const T inv_sqrt_2pi = (T)0.39894228040143267794;
  const T inv_sqrt_pi = (T)0.56418958354775628695;
  T* dst = nodal + b * (int64_t)N * J;
  const int M = (Rrows + 1) / 2;
  for (int idx = tid; idx < N * TJ; idx += nt) {
    int n = idx / TJ, c = idx % TJ, j = j0 + c;
    if (j >= J) continue;
    T acc = g[c] * inv_sqrt_2pi;
    T accm = 0;
    int k = 0;
    for (int m = 1; m < M; ++m) {
      k += n;
      if (k >= N) k -= N;
      accm += g[(2 * m - 1) * TJ + c] * cs[2 * k] + g[(2 * m) * TJ + c] * cs[2 * k + 1];
    }
    T v = acc + accm * inv_sqrt_pi;
    if (weighted) v *= qw[j];
    dst[(int64_t)n * J + j] = v;
  }
}


template <typename T>
int fft_pairs_for(int N) {
  // 2P columns per workgroup: 32 columns (128 B rows in FP32) while the LDS
  // footprint N * P * sizeof(complex) stays <= 64 KiB.
  int P = 16;
  while (P > 1 && (size_t)N * P * 2 * sizeof(T) > 65536) P /= 2;
  return P;
}

template <typename T, int N, int P>
int launch_fft_fixed(bool forward, const dino_plan* pl, const T* in, T* out,
                     const T* tw, const T* qw, int64_t B, int weighted, hipStream_t s) {
  dim3 grid((pl->J + 2 * P - 1) / (2 * P), (unsigned)B);
  size_t lds = sizeof(cplx<T>) * (size_t)N * (P + 1);
  if (forward) {
    auto k = dft_forward_fft_kernel<T, N, P, kFftThreads>;
    if (lds > 65536) DINO_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k<<<grid, kFftThreads, lds, s>>>(in, out, tw, qw, pl->J, pl->R, weighted);
  } else {
    auto k = dft_inverse_fft_kernel<T, N, P, kFftThreads>;
    if (lds > 65536) DINO_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k<<<grid, kFftThreads, lds, s>>>(in, out, tw, qw, pl->J, pl->M, weighted);
  }
  DINO_LAUNCH_CHECK("dft fft kernel");
  return DINO_OK;
}

int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
      n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

template <int N>
int launch_fft_persistent(bool forward, const dino_plan* pl, const float* in, float* out,
                          const float* tw, const float* qw, int64_t B, int weighted,
                          hipStream_t s) {
  const int tpf = pl->J / 32;
  const int64_t ntiles = B * tpf;
  const size_t lds = sizeof(cplx<float>) * (size_t)N * 17;
  // one resident wave of workgroups: as many per CU as registers and LDS allow
  auto kf = fft_forward_persistent_kernel<N>;
  auto ki = fft_inverse_persistent_kernel<N>;
  const void* k = forward ? (const void*)kf : (const void*)ki;
  if (lds > 65536)
    DINO_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  DINO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kPersistThreads, lds));
  const int64_t grid = std::min<int64_t>(ntiles, (int64_t)std::max(per_cu, 1) * device_cus());
  if (forward)
    kf<<<(unsigned)grid, kPersistThreads, lds, s>>>(in, out, tw, qw, pl->J, pl->R, weighted, ntiles, tpf);
  else
    ki<<<(unsigned)grid, kPersistThreads, lds, s>>>(in, out, tw, qw, pl->J, pl->M, weighted, ntiles, tpf);
  DINO_LAUNCH_CHECK("persistent fft kernel");
  return DINO_OK;
}

template <typename T, int N>
int launch_fft_n(bool forward, const dino_plan* pl, const T* in, T* out, const T* tw,
                 const T* qw, int64_t B, int weighted, hipStream_t s) {
  if constexpr (sizeof(T) == 4 && N >= 128 && N <= 512) {
    const bool aligned = reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                         reinterpret_cast<uintptr_t>(out) % 16 == 0;
    const int fam = pl->sched[forward ? DINO_STAGE_DFT_FORWARD : DINO_STAGE_DFT_INVERSE] & 0xff;
    if (fam != DINO_SCHED_PER_TILE && pl->J % 32 == 0 && aligned && 2 * pl->M - 1 <= N)
      return launch_fft_persistent<N>(forward, pl, in, out, tw, qw, B, weighted, s);
  }
  int P = fft_pairs_for<T>(N);
  switch (P) {
    case 16: return launch_fft_fixed<T, N, 16>(forward, pl, in, out, tw, qw, B, weighted, s);
    case 8: return launch_fft_fixed<T, N, 8>(forward, pl, in, out, tw, qw, B, weighted, s);
    case 4: return launch_fft_fixed<T, N, 4>(forward, pl, in, out, tw, qw, B, weighted, s);
    default: return launch_fft_fixed<T, N, 2>(forward, pl, in, out, tw, qw, B, weighted, s);
  }
}

template <typename T>
int run_dft(bool forward, const dino_plan* pl, const void* in_v, void* out_v, int64_t B,
            int weighted, hipStream_t s) {
  const T* in = static_cast<const T*>(in_v);
  T* out = static_cast<T*>(out_v);
  const T* tw;
  const T* qw;
  if constexpr (sizeof(T) == 8) {
    tw = pl->tw64;
    qw = pl->qw64;
  } else {
    tw = pl->tw32;
    qw = pl->qw32;
  }
  if (B == 0) return DINO_OK;
  DINO_REQUIRE(B <= 65535, DINO_EINVAL, "batch %lld exceeds 65535 fields per call",
               (long long)B);
  if (pl->pow2) {
    switch (pl->I) {
      case 8: return launch_fft_n<T, 8>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 16: return launch_fft_n<T, 16>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 32: return launch_fft_n<T, 32>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 64: return launch_fft_n<T, 64>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 128: return launch_fft_n<T, 128>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 256: return launch_fft_n<T, 256>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 512: return launch_fft_n<T, 512>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 1024: return launch_fft_n<T, 1024>(forward, pl, in, out, tw, qw, B, weighted, s);
      case 2048: return launch_fft_n<T, 2048>(forward, pl, in, out, tw, qw, B, weighted, s);
      default: break;  // 4096 falls through to the direct path
    }
  }
  int N = pl->I;
  int rows = forward ? N : pl->R;
  int TJ = 32;
  while (TJ > 1 && (size_t)(rows * TJ + 2 * N) * sizeof(T) > 65536) TJ /= 2;
  size_t lds = (size_t)(rows * TJ + 2 * N) * sizeof(T);
  DINO_REQUIRE(lds <= 65536, DINO_EINVAL,
               "longitude transform of %d nodes does not fit the direct DFT path", N);
  dim3 grid((pl->J + TJ - 1) / TJ, (unsigned)B);
  if (forward)
    dft_forward_direct_kernel<T><<<grid, 256, lds, s>>>(in, out, tw, qw, N, pl->J, pl->R, TJ, weighted);
  else
    dft_inverse_direct_kernel<T><<<grid, 256, lds, s>>>(in, out, tw, qw, N, pl->J, pl->R, TJ, weighted);
  DINO_LAUNCH_CHECK("dft direct kernel");
  return DINO_OK;
}

}  // namespace

int dft_forward_impl(const dino_plan* pl, const void* nodal, void* fourier, int64_t B,
                     int dtype, int weighted, hipStream_t s) {
  if (dtype == DINO_F64) return run_dft<double>(true, pl, nodal, fourier, B, weighted, s);
  return run_dft<float>(true, pl, nodal, fourier, B, weighted, s);
}

int dft_inverse_impl(const dino_plan* pl, const void* fourier, void* nodal, int64_t B,
                     int dtype, int weighted, hipStream_t s) {
  if (dtype == DINO_F64) return run_dft<double>(false, pl, fourier, nodal, B, weighted, s);
  return run_dft<float>(false, pl, fourier, nodal, B, weighted, s);
}

}  // namespace dino
