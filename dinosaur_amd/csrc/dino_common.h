// Internal definitions shared by the dinosaur_amd HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/dinosaur_amd.h"

namespace dino {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define DINO_HIP(call)                                   \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return ::dino::hip_fail(e_, #call); \
  } while (0)

#define DINO_LAUNCH_CHECK(what)                          \
  do {                                                   \
    hipError_t e_ = hipGetLastError();                   \
    if (e_ != hipSuccess) return ::dino::hip_fail(e_, what); \
  } while (0)

#define DINO_REQUIRE(cond, code, ...)                    \
  do {                                                   \
    if (!(cond)) {                                       \
      ::dino::set_error(__VA_ARGS__);                    \
      return code;                                       \
    }                                                    \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace dino

// Transform plan: everything lives on the device and is immutable after
// creation, so a plan may be shared by threads and streams.
struct dino_plan {
  int M, L, I, J;
  int R;        // 2M - 1 modal rows
  int Jh;       // ceil(J / 2): rows of the folded (northern) hemisphere
  int Jhp;      // Jh padded to a multiple of 8 (table row stride)
  bool pow2;    // I is a power of two -> FFT path
  int64_t table_elems;  // sum_m (L - m) * Jhp
  // device arrays
  double* nodes64;      // (J) ascending
  double* weights64;    // (J) Gauss weights (sum 2)
  float* qw32;          // (J) (2 pi / I) * w_j
  double* qw64;
  int64_t* m_offset;    // (M) start of row l = m of each m block in the table
  double* p64;          // folded table, FP64
  float* p32;           // FP32 copy
  double* tw64;         // (I) complex twiddles exp(-2 pi i k / I) as (cos, sin)
  float* tw32;
  // Forward-transform tables in MFMA fragment order (legendre_rows.hip).
  // fks = number of 32-node reduction steps the tables hold (1, 2, 4 or 8;
  // 0 when Jh > 256 and the tiled kernels are used instead).
  int fks;
  int64_t* ftile_off;   // (M) first tile of each m, in 16-byte units
  float* ffrag32;       // FP32 fragments
  uint16_t* ffragbf;    // BF16 hi/lo fragments
  float* sink;          // (256) store target of out-of-range lanes, never read
  // Inverse-transform tables in MFMA fragment order (legendre_inv_rows.hip).
  // int_nt = 16-node tiles the tables hold (1, 2, 4 or 8; 0 when Jh > 128).
  int int_nt;
  int64_t* itile_off;   // (M) first tile of each m, in 16-byte units
  float* ifrag32;       // FP32 fragments
  float* ifragbf;       // BF16 hi/lo fragments (16-byte slots)
  // Launch schedule per stage (DINO_STAGE_*): DINO_SCHED_* family in the low
  // byte, DINO_SCHED_MPAIRED flag.  Zero (auto) after creation.
  int sched[4];
};

// Degree tiles of the row-resident Legendre kernels: 32-degree tiles that
// cover l = m .. L-1, aligned so that the last one ends at L -- the top
// degree then never sits alone in a tile -- unless the first would start
// below degree 0 (then the first starts at 0 and the last runs past L).
constexpr int kLegTile = 32;
__host__ __device__ inline int leg_tile_start(int m, int L) {
  const int s = L - kLegTile * ((L - m + kLegTile - 1) / kLegTile);
  return s > 0 ? s : 0;
}
__host__ __device__ inline int leg_tile_count(int m, int L) {
  return (L - leg_tile_start(m, L) + kLegTile - 1) / kLegTile;
}

// Orders this wave's LDS writes before its later LDS reads when the staging
// region is private to the wave: no workgroup barrier, so the waves of a
// block need not wait for each other.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups of a launch are dispatched to the GPU's 8 XCDs, each with its
// own L2, in turn: hardware workgroup n (x fastest) runs on XCD n mod 8.
// xcd_block maps it to a logical workgroup (bx, by) such that each XCD takes
// one contiguous eighth of the logical grid, so workgroups that read each
// other's neighbours (stencils across rows) share an L2.  A permutation of
// the grid: correct whatever the dispatch order is.
__device__ __forceinline__ void xcd_block(int* bx, int* by) {
  const int gx = (int)gridDim.x;
  const int64_t n = (int64_t)gx * gridDim.y;
  const int64_t lin = blockIdx.x + (int64_t)blockIdx.y * gx;
  const int64_t q = n / 8, r = n % 8, k = lin % 8, i = lin / 8;
  const int64_t logical = (k < r ? k * (q + 1) : r * (q + 1) + (k - r) * q) + i;
  *bx = (int)(logical % gx);
  *by = (int)(logical / gx);
}

// Zonal wavenumber handled by grid row y of a Legendre launch over M rows:
// natural order (longest triangular slab first) or paired (0, M-1, 1, M-2,
// ...), which alternates long and short slabs.
__device__ __forceinline__ int m_of_block(int y, int M, int paired) {
  return paired ? ((y & 1) ? M - 1 - (y >> 1) : (y >> 1)) : y;
}

struct dino_implicit {
  int K, L, R;
  double r_gas, eta;
  // device FP64: G, H, V, Vinv (K*K each), tref, dsigma, d (K each), lam (L)
  double* G;
  double* H;
  double* V;
  double* Vinv;
  double* tref;
  double* dsigma;
  double* d;
  double* lam;
  // fragment tables and vectors of the MFMA kernels, forward [0..7] and
  // transposed [8..14] (layout in implicit_mfma.hip)
  double* mtab[15];
};
