// Forward Legendre contraction with register-resident rows.
//
// D[row, l] = sum_jj X_{(l-m) mod 2}[row, jj] P[l, jj], X_0 = f_up + f_dn,
// X_1 = f_up - f_dn (equatorial fold; same contraction as legendre.hip).
//
// The contraction is HBM-bound at its batch sizes (one read of the Fourier
// rows, one write of the modal rows, including the zero block l < m), so
// the kernel is laid out to touch each byte of HBM once:
//   * a workgroup owns 64 rows (4 waves x 16) of one zonal wavenumber m and
//     sweeps every degree tile of that m, so the Fourier rows are read once;
//   * each wave keeps the folded rows of its 16-row strip in VGPRs as MFMA
//     A-operand fragments for the whole reduction (all Jh northern nodes),
//     split once into BF16 hi/lo for the split-precision path;
//   * the Legendre values arrive through a double-buffered LDS tile that is
//     filled from a plan table stored in fragment order (16-byte lane reads,
//     conflict-free) -- build_forward_fragments below.
//
// Reduction order inside a 16-node group t: lane group g = lane / 16 holds
// nodes jj = 16 t + 4 g + e (e = 0..3), one float4 of the row per lane.  A
// degree tile covers 32 consecutive degrees lt .. lt + 31; the tiles of one m
// end at the last degree (lt = leg_tile_start(m, L) + 32 k), so the zero
// degrees l < m fill at most 31 slots of the first tile.  Accumulator q (0 or
// 1) holds the degrees lt + 2 c + q, c = lane % 16.
#include "dino_common.h"

namespace dino {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kRowsPerWave = 16;
constexpr int kTileDeg = 32;
constexpr int kRingPitch = 68;  // output ring row pitch (floats): 64 degrees + 4
static_assert(kRingPitch % 4 == 0 && kRingPitch >= 64, "float4-aligned rows of 64 degrees");

__device__ __forceinline__ void row_field(int m, int q, int* b, int* r) {
  if (m == 0) {
    *b = q;
    *r = 0;
  } else {
    *b = q >> 1;
    *r = 2 * m - 1 + (q & 1);
  }
}

static_assert(kTileDeg == kLegTile, "tile width shared with the modal kernels");
__host__ __device__ inline int tiles_of_m(int m, int L) { return leg_tile_count(m, L); }

// float4 units per degree tile: 2 parities x 2 KS node groups x 64 lanes for
// FP32, 2 parities x KS steps x (hi, lo) x 64 lanes for BF16 -- both 256 KS.
__host__ __device__ constexpr int tile_f4(int ks) { return 256 * ks; }

__device__ __forceinline__ bf16x8 split_hi(const float4& a, const float4& b) {
  bf16x8 h;
  h[0] = (__bf16)a.x; h[1] = (__bf16)a.y; h[2] = (__bf16)a.z; h[3] = (__bf16)a.w;
  h[4] = (__bf16)b.x; h[5] = (__bf16)b.y; h[6] = (__bf16)b.z; h[7] = (__bf16)b.w;
  return h;
}

__device__ __forceinline__ bf16x8 split_lo(const float4& a, const float4& b, const bf16x8& h) {
  bf16x8 l;
  l[0] = (__bf16)(a.x - (float)h[0]); l[1] = (__bf16)(a.y - (float)h[1]);
  l[2] = (__bf16)(a.z - (float)h[2]); l[3] = (__bf16)(a.w - (float)h[3]);
  l[4] = (__bf16)(b.x - (float)h[4]); l[5] = (__bf16)(b.y - (float)h[5]);
  l[6] = (__bf16)(b.z - (float)h[6]); l[7] = (__bf16)(b.w - (float)h[7]);
  return l;
}

__device__ __forceinline__ bf16x8 as_bf16x8(const floatx4& v) {
  return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ floatx4 ld4(const float4* p) {
  return *reinterpret_cast<const floatx4*>(p);
}

// Float index mod 32 of an address (its position in a 128-byte line).
__device__ __forceinline__ int line_pos(const float* p) {
  return (int)((reinterpret_cast<uintptr_t>(p) >> 2) & 31);
}

// Zero-block extent of a row of zonal wavenumber m starting at p: the
// degrees [0, n) below the 128-byte line that holds degree lt0 (the tile loop
// writes that line, zeros included); n < 0 when that line starts before p.
__device__ __forceinline__ int zero_extent(const float* p, int m, int L) {
  const int lt0 = leg_tile_start(m, L);
  return lt0 - ((line_pos(p) + lt0) & 31);
}

// Number of floats of row (b, r + 1) that the writer of row (b, r) stores
// as zeros: the head of row r + 1 up to its first line boundary, when that
// head lies in row r + 1's zero block (the line shared by the two rows is
// then written whole, by one writer); 0 otherwise.  next = start of row r + 1.
__device__ __forceinline__ int next_row_head(const float* next, int r, int Rrows, int L) {
  if (r + 1 >= Rrows) return 0;
  const int s = (32 - line_pos(next)) & 31;
  return zero_extent(next, (r + 2) >> 1, L) >= 0 ? s : 0;
}

// Zero block of the 16 rows strip0 .. strip0 + 15 (one wave), skipped on
// request: degrees [s, n) of each row, n = zero_extent, s = the row's first
// line boundary (the head [0, s) is stored by the previous row's writer, see
// next_row_head), so only whole 128-byte lines.  One row per step with
// uniform (scalar) row addresses.
__device__ __forceinline__ void zero_block(float* __restrict__ modal, int m, int strip0,
                                           int nrows, int Rrows, int L, int lt0, int lane,
                                           int skip_zero) {
  if (lt0 == 0 || skip_zero) return;
  const int s0 = __builtin_amdgcn_readfirstlane(strip0);
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < 16 && s0 + i < nrows; ++i) {
    int b, r;
    row_field(m, s0 + i, &b, &r);
    float* const p = modal + ((int64_t)b * Rrows + r) * L;
    const int n = zero_extent(p, m, L);
    const int s = (32 - line_pos(p)) & 31;
    for (int c4 = lane; 4 * c4 < n - s; c4 += 64)
      *reinterpret_cast<floatx4*>(p + s + 4 * c4) = zero;
  }
}

// VIN: J % 8 == 0 and a 16-byte aligned Fourier buffer (coalesced float4 row
// loads staged through LDS), otherwise a scalar gather for any J.  The modal
// rows may have any length and alignment (TL grids have odd L).
// NW waves (16 rows each) per workgroup.  Every workgroup streams all P tiles
// of its m from L2, so 8 waves (128 rows) halve that traffic per row
// against 4 -- these requests queue with the HBM reads for the same L1
// miss slots.  Up to 128 northern nodes (KS <= 4) the registers allow 3
// waves per SIMD with 4-wave workgroups (3 per CU, as the LDS budget does)
// and 4 per SIMD with 8-wave ones (2 per CU).
template <int NW, int KS>
constexpr int fwd_waves_per_simd() {
  return KS > 4 ? 1 : (NW == 8 ? 4 : 3);
}

template <bool BF, int KS, bool VIN, int NW>
__global__ __launch_bounds__(64 * NW, (fwd_waves_per_simd<NW, KS>())) void leg_fwd_rows_kernel(
    const float* __restrict__ fourier, float* __restrict__ modal,
    const float4* __restrict__ ftab, const int64_t* __restrict__ ftile_off,
    float* __restrict__ sink, int B, int L, int J, int Jh, int Rrows, int mpair, int skip_zero) {
  constexpr int KT = 2 * KS;  // 16-node groups
  constexpr int NTH = 64 * NW;  // threads
  constexpr int RPB = kRowsPerWave * NW;  // rows per workgroup
  constexpr int TF4 = tile_f4(KS);
  constexpr int PER = (TF4 + NTH - 1) / NTH;
  // One LDS region: first the staging of the folded rows (RPB rows, pitch
  // padded by 4 floats against bank conflicts), then the table double buffer
  // followed by the output rings (one per wave).
  constexpr int PITCH = 32 * KS + 4;
  constexpr int STG4 = RPB * PITCH / 4;
  constexpr int RING4 = NW * kRowsPerWave * kRingPitch / 4;
  constexpr int LDS4 = (2 * TF4 + RING4 > STG4) ? 2 * TF4 + RING4 : STG4;
  __shared__ floatx4 lds[LDS4];
  float* const ring = reinterpret_cast<float*>(lds + 2 * TF4);
  const int m = m_of_block(blockIdx.y, gridDim.y, mpair);
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.x * RPB;
  if (row0 >= nrows) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c16 = lane & 15;
  const int lt0 = leg_tile_start(m, L);
  const int ntile = tiles_of_m(m, L);
  const float4* tb = ftab + ftile_off[m];

  // Table tile prefetch (clang vector type, unconditional loads: both keep
  // the staging array in VGPRs so the loads stay asynchronous).
  floatx4 st[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int idx = tid + NTH * i;
    st[i] = ld4(tb + (idx < TF4 ? idx : 0));
  }

  // Folded rows: X0 pairs with the even-d degrees of a tile (l - m of
  // parity (lt0 - m) & 1), X1 with the odd-d ones.  When lt0 - m is odd the
  // roles of the sum and the difference swap.
  const float sgn = ((lt0 - m) & 1) ? -1.f : 1.f;
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  float4 x0[KT], x1[KT];
  if constexpr (VIN) {
    // Coalesced loads: lane = (row rho of the wave's strip, 4-node chunk c)
    // with CPR chunks per half row (the whole reduction width) and RPI rows
    // per instruction, both compile-time, so no index division; northern
    // half ascending, its mirror image descending.  Chunks past Jh / 4 and
    // rows past the batch read a clamped in-range address and are zeroed by
    // selects, so the loads issue back to back without branches.
    constexpr int CPR = 8 * KS, RPI = 64 / CPR;
    static_assert(KT * RPI == 16, "KT load instructions cover the wave's 16 rows");
    const int H4 = Jh / 4;
    const int c = lane % CPR, cc = c < H4 ? c : H4 - 1;
    floatx4 u[KT], d[KT];
#pragma unroll
    for (int it = 0; it < KT; ++it) {
      const int row = row0 + wave * 16 + lane / CPR + RPI * it;
      const bool ok = c < H4 && row < nrows;
      int b, r;
      row_field(m, row < nrows ? row : nrows - 1, &b, &r);
      const float* f = fourier + ((int64_t)b * Rrows + r) * J;
      const floatx4 uu = *reinterpret_cast<const floatx4*>(f + J - Jh + 4 * cc);
      const floatx4 dr = *reinterpret_cast<const floatx4*>(f + Jh - 4 - 4 * cc);
      u[it] = ok ? uu : zero;
      d[it] = ok ? floatx4{dr[3], dr[2], dr[1], dr[0]} : zero;
    }
    // The zero block is written while these loads are in flight.
    zero_block(modal, m, row0 + wave * 16, nrows, Rrows, L, lt0, lane, skip_zero);
    // Transpose through LDS into A-operand fragments, X0 first, then X1.
    float* stg = reinterpret_cast<float*>(lds);
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const float sp = pass == 0 ? sgn : -sgn;
#pragma unroll
      for (int it = 0; it < KT; ++it) {
        const int rho = lane / CPR + RPI * it;
        *reinterpret_cast<floatx4*>(stg + (wave * 16 + rho) * PITCH + 4 * c) = u[it] + sp * d[it];
      }
      wave_lds_sync();  // each wave stages only its own 16 rows
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        const int jj = 16 * t + 4 * g;
        const floatx4 v = *reinterpret_cast<const floatx4*>(stg + (wave * 16 + c16) * PITCH + jj);
        const floatx4 w = jj < Jh ? v : zero;
        if (pass == 0) x0[t] = make_float4(w[0], w[1], w[2], w[3]);
        else x1[t] = make_float4(w[0], w[1], w[2], w[3]);
      }
      // pass 1: the whole block reuses the staging LDS as table buffers and
      // rings next
      if (pass == 0) wave_lds_sync();
      else __syncthreads();
    }
  } else {
    // Generic shapes: gather the fragments directly.
    const int row = row0 + wave * 16 + c16;
    const bool rok = row < nrows;
    int b = 0, r = 0;
    if (rok) row_field(m, row, &b, &r);
    const float* f = fourier + ((int64_t)b * Rrows + r) * J;
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      const int jj = 16 * t + 4 * g;
      floatx4 u, d;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int q = jj + e;
        const bool ok = rok && q < Jh;
        // odd J: node jj = 0 is the equator, counted once
        const bool okd = ok && !((J & 1) && q == 0);
        const float uu = f[ok ? J - Jh + q : 0];
        const float dd = f[okd ? Jh - 1 - q : 0];
        u[e] = ok ? uu : 0.f;
        d[e] = okd ? dd : 0.f;
      }
      x0[t] = make_float4(u[0] + sgn * d[0], u[1] + sgn * d[1], u[2] + sgn * d[2],
                          u[3] + sgn * d[3]);
      x1[t] = make_float4(u[0] - sgn * d[0], u[1] - sgn * d[1], u[2] - sgn * d[2],
                          u[3] - sgn * d[3]);
    }
    zero_block(modal, m, row0 + wave * 16, nrows, Rrows, L, lt0, lane, skip_zero);
  }
  bf16x8 h0[BF ? KS : 1], l0[BF ? KS : 1], h1[BF ? KS : 1], l1[BF ? KS : 1];
  if constexpr (BF) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      h0[s] = split_hi(x0[2 * s], x0[2 * s + 1]);
      l0[s] = split_lo(x0[2 * s], x0[2 * s + 1], h0[s]);
      h1[s] = split_hi(x1[2 * s], x1[2 * s + 1]);
      l1[s] = split_lo(x1[2 * s], x1[2 * s + 1], h1[s]);
    }
  }

  // Output.  Every row is written in whole 128-byte lines aligned in memory
  // (float4 stores): a TL row holds an odd number of degrees, so tile-sized
  // segments would start anywhere in a line and split into partial writes.
  // A per-row LDS ring of 64 degrees (two tiles) keeps the degrees of the
  // previous tile that a line straddling the tile boundary still needs;
  // degree l of a row sits at ring slot (ra + l) & 63, ra = row address / 4
  // mod 64, so a line is contiguous in the ring.  With a = (ra + lt0) & 31,
  // tile lt writes the line starting at degree lt - a (and the next one on
  // the last tile).  The degrees lt0 - a .. lt0 - 1 of the first line are
  // zeros from the ring, also in triangle-only mode (entries below lt0 may be
  // zeroed there), and a row's last line also takes the next row's zero head
  // (next_row_head), so lines are whole except where a row's first line
  // begins in the previous row's non-zero tail: quads with degrees outside
  // [0, hi) are written float by float.  Out-of-range lanes store into
  // `sink` instead of branching around the store, so every store
  // instruction issues and the next tile's wait for its table loads can
  // leave the stores in flight.
  float* const obase = modal + (int64_t)(m == 0 ? row0 : row0 >> 1) * Rrows * L;
  auto row_offset = [&](int rl) -> int {  // from obase (at most 128 fields: 32 bits), -1 if none
    const int row = row0 + wave * 16 + rl;
    if (row >= nrows) return -1;
    int b, r;
    row_field(m, row, &b, &r);
    return ((b - (m == 0 ? row0 : row0 >> 1)) * Rrows + r) * L;
  };
  auto ring_origin = [&](int off) -> int {
    return off < 0 ? 0 : (int)((reinterpret_cast<uintptr_t>(obase + off) >> 2) & 63);
  };
  // rows this lane stores: quad store j -> row 8 j + lane / 8, quad lane % 8;
  // hi = end of what the row's writer stores (L plus the next row's head)
  int orow[2], ra_s[2], hi[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    orow[j] = row_offset(8 * j + (lane >> 3));
    ra_s[j] = ring_origin(orow[j]);
    hi[j] = L;
    if (orow[j] >= 0) {
      int b, r;
      row_field(m, row0 + wave * 16 + 8 * j + (lane >> 3), &b, &r);
      hi[j] = L + next_row_head(obase + orow[j] + L, r, Rrows, L);
    }
  }
  // rows this lane fills in the ring: 4 g + i
  int ra_w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) ra_w[i] = ring_origin(row_offset(4 * g + i));
  float* const rg = ring + wave * 16 * kRingPitch;
  // the ring starts as zeros: the degrees just below lt0 that the first
  // tile's first line covers are zero-block degrees
  for (int i = lane; i < 16 * kRingPitch / 4; i += 64) reinterpret_cast<floatx4*>(rg)[i] = zero;

  for (int tt = 0; tt < ntile; ++tt) {
    floatx4* buf = lds + (tt & 1) * TF4;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = tid + NTH * i;
      if (idx < TF4) buf[idx] = st[i];
    }
    __syncthreads();
    {
      // the last iteration reloads its own tile, which is never stored
      const float4* nb = tb + (int64_t)(tt + 1 < ntile ? tt + 1 : tt) * TF4;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int idx = tid + NTH * i;
        st[i] = ld4(nb + (idx < TF4 ? idx : 0));
      }
    }
    // keep the prefetch ahead of the MFMAs (the scheduler would otherwise
    // sink the loads to the end of the tile, serialising them with it)
    __builtin_amdgcn_sched_barrier(0);
    floatx4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = a0;
    if constexpr (BF) {
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const bf16x8 bh0 = as_bf16x8(buf[((0 * KS + s) * 2 + 0) * 64 + lane]);
        const bf16x8 bl0 = as_bf16x8(buf[((0 * KS + s) * 2 + 1) * 64 + lane]);
        const bf16x8 bh1 = as_bf16x8(buf[((1 * KS + s) * 2 + 0) * 64 + lane]);
        const bf16x8 bl1 = as_bf16x8(buf[((1 * KS + s) * 2 + 1) * 64 + lane]);
        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0[s], bh0, a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l1[s], bh1, a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0[s], bl0, a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1[s], bl1, a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0[s], bh0, a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1[s], bh1, a1, 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        const floatx4 p0 = buf[(0 * KT + t) * 64 + lane];
        const floatx4 p1 = buf[(1 * KT + t) * 64 + lane];
        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[t].x, p0[0], a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[t].x, p1[0], a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[t].y, p0[1], a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[t].y, p1[1], a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[t].z, p0[2], a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[t].z, p1[2], a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[t].w, p0[3], a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[t].w, p1[3], a1, 0, 0, 0);
      }
    }
    const int lt = lt0 + tt * kTileDeg;
    // this tile's degrees lt + 2 c16 + q of rows 4 g + i into the ring
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float* const rr = rg + (4 * g + i) * kRingPitch;
      rr[(ra_w[i] + lt + 2 * c16) & 63] = a0[i];
      rr[(ra_w[i] + lt + 2 * c16 + 1) & 63] = a1[i];
    }
    wave_lds_sync();  // the ring is private to the wave
    const int nk = tt + 1 < ntile ? 1 : 2;
#pragma unroll
    for (int jk = 0; jk < 4; ++jk) {
      const int j = jk & 1, k = jk >> 1;
      if (k >= nk) break;
      const int rl = 8 * j + (lane >> 3);
      const int l4 = lt - ((ra_s[j] + lt0) & 31) + 32 * k + 4 * (lane & 7);
      const floatx4 v =
          *reinterpret_cast<const floatx4*>(rg + rl * kRingPitch + ((ra_s[j] + l4) & 63));
      floatx4 w;  // degrees >= L: the next row's zero head
#pragma unroll
      for (int e = 0; e < 4; ++e) w[e] = l4 + e < L ? v[e] : 0.f;
      const bool rok = orow[j] >= 0;
      const bool full = rok && l4 >= 0 && l4 + 3 < hi[j];
      floatx4* const dst = full ? reinterpret_cast<floatx4*>(obase + orow[j] + l4)
                                : reinterpret_cast<floatx4*>(sink) + lane;
      *dst = w;
      if (rok && !full && l4 + 3 >= 0 && l4 < hi[j]) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (l4 + e >= 0 && l4 + e < hi[j]) obase[orow[j] + l4 + e] = w[e];
      }
    }
    wave_lds_sync();  // reads of this tile's lines before the next tile's ring writes
  }
}

// One thread per fragment slot of every tile of one m (blockIdx.y = m).
// FP32 slot order [tile][q][t][lane][e]; BF16 [tile][q][s][h][lane][e'].
__global__ void forward_fragments_kernel(int L, int Jh, int Jhp, int ks,
                                         const double* __restrict__ p64,
                                         const int64_t* __restrict__ m_off,
                                         const int64_t* __restrict__ ftile_off,
                                         float* __restrict__ f32, uint16_t* __restrict__ fbf) {
  const int m = blockIdx.y;
  const int64_t per_tile = 1024 * (int64_t)ks;  // slots (floats) per tile
  const int64_t n = per_tile * tiles_of_m(m, L);
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int tile = (int)(idx / per_tile);
  const int rem = (int)(idx - tile * per_tile);
  const int lt = leg_tile_start(m, L) + kTileDeg * tile;
  auto value = [&](int l, int jj) -> double {
    return (l >= m && l < L && jj < Jh) ? p64[m_off[m] + (int64_t)(l - m) * Jhp + jj] : 0.0;
  };
  const int64_t base = ftile_off[m] * 4 + (int64_t)tile * per_tile;  // floats
  {
    // FP32: rem = ((q * KT + t) * 64 + lane) * 4 + e
    const int KT = 2 * ks;
    const int q = rem / (KT * 256), r2 = rem % (KT * 256);
    const int t = r2 / 256, lane = (r2 % 256) / 4, e = r2 % 4;
    const int l = lt + 2 * (lane % 16) + q;
    const int jj = 16 * t + 4 * (lane / 16) + e;
    f32[base + rem] = (float)value(l, jj);
  }
  {
    // BF16: rem = ((q * KS + s) * 64 + lane) * 8 + e'; hi at h = 0, lo at h = 1
    const int q = rem / (ks * 512), r2 = rem % (ks * 512);
    const int s = r2 / 512, lane = (r2 % 512) / 8, e = r2 % 8;
    const int l = lt + 2 * (lane % 16) + q;
    const int jj = 16 * (2 * s + e / 4) + 4 * (lane / 16) + e % 4;
    const double v = value(l, jj);
    const __bf16 hi = (__bf16)(float)v;
    const __bf16 lo = (__bf16)(float)(v - (double)(float)hi);
    const int64_t ub = base * 2;  // uint16 units
    const int64_t pos = ((int64_t)(q * ks + s) * 2) * 512 + lane * 8 + e;
    fbf[ub + pos] = __builtin_bit_cast(uint16_t, hi);
    fbf[ub + pos + 512] = __builtin_bit_cast(uint16_t, lo);
  }
}

template <bool BF, int KS, int NW>
int launch_rows(const dino_plan* pl, const float* in, float* out, int64_t B, int skip_zero,
                hipStream_t s) {
  constexpr int RPB = kRowsPerWave * NW;
  const bool vin = pl->J % 8 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0;
  const dim3 grid((unsigned)((2 * B + RPB - 1) / RPB), (unsigned)pl->M);
  const float4* tab = BF ? reinterpret_cast<const float4*>(pl->ffragbf)
                         : reinterpret_cast<const float4*>(pl->ffrag32);
  const int mpair = (pl->sched[DINO_STAGE_LEGENDRE_FORWARD] & DINO_SCHED_MPAIRED) ? 1 : 0;
#define ROWS_LAUNCH(VI)                                                                   \
  leg_fwd_rows_kernel<BF, KS, VI, NW><<<grid, 64 * NW, 0, s>>>(                            \
      in, out, tab, pl->ftile_off, pl->sink, (int)B, pl->L, pl->J, pl->Jh, pl->R, mpair,  \
      skip_zero)
  if (vin) ROWS_LAUNCH(true);
  else ROWS_LAUNCH(false);
#undef ROWS_LAUNCH
  DINO_LAUNCH_CHECK("legendre forward rows kernel");
  return DINO_OK;
}

template <bool BF, int NW>
int rows_by_ks(const dino_plan* pl, const float* in, float* out, int64_t B, int skip_zero,
               hipStream_t s) {
  switch (pl->fks) {
    case 1: return launch_rows<BF, 1, NW>(pl, in, out, B, skip_zero, s);
    case 2: return launch_rows<BF, 2, NW>(pl, in, out, B, skip_zero, s);
    case 4: return launch_rows<BF, 4, NW>(pl, in, out, B, skip_zero, s);
    case 8: return launch_rows<BF, 8, NW>(pl, in, out, B, skip_zero, s);
    default:
      set_error("forward fragment tables hold %d steps", pl->fks);
      return DINO_EINVAL;
  }
}

}  // namespace

// Number of 32-node steps the fragment tables use for this plan (a power of
// two covering Jh), or 0 when Jh is too large for register-resident rows.
int forward_fragment_steps(int Jh) {
  int ks = 1;
  while (32 * ks < Jh) ks *= 2;
  return ks <= 8 ? ks : 0;
}

// Float4 units of the fragment tables; fills off[m] (float4 units).
int64_t forward_fragment_layout(int M, int L, int ks, int64_t* off) {
  int64_t acc = 0;
  for (int m = 0; m < M; ++m) {
    off[m] = acc;
    acc += (int64_t)tiles_of_m(m, L) * tile_f4(ks);
  }
  return acc;
}

void forward_fragment_fill(const dino_plan* pl, hipStream_t s) {
  const int64_t most = 1024 * (int64_t)pl->fks * tiles_of_m(0, pl->L);
  const dim3 grid((unsigned)((most + 255) / 256), (unsigned)pl->M);
  forward_fragments_kernel<<<grid, 256, 0, s>>>(pl->L, pl->Jh, pl->Jhp, pl->fks, pl->p64,
                                                pl->m_offset, pl->ftile_off, pl->ffrag32,
                                                pl->ffragbf);
}

bool forward_rows_available(const dino_plan* pl) { return pl->fks > 0; }

int legendre_forward_rows(const dino_plan* pl, const float* in, float* out, int64_t B,
                          bool split, bool triangle_only, bool wide, hipStream_t s) {
  if (B == 0) return DINO_OK;
  DINO_REQUIRE(2 * B / kRowsPerWave < (int64_t)1 << 30, DINO_EINVAL, "batch too large");
  // a block's output rows span at most 128 fields, addressed with 32-bit offsets
  DINO_REQUIRE((int64_t)8 * kRowsPerWave * pl->R * pl->L < (int64_t)1 << 31, DINO_EINVAL,
               "modal fields too large for the row kernel");
  const int z = triangle_only ? 1 : 0;
  if (wide)
    return split ? rows_by_ks<true, 8>(pl, in, out, B, z, s)
                 : rows_by_ks<false, 8>(pl, in, out, B, z, s);
  return split ? rows_by_ks<true, 4>(pl, in, out, B, z, s)
               : rows_by_ks<false, 4>(pl, in, out, B, z, s);
}

}  // namespace dino
