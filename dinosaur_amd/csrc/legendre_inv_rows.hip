// Inverse Legendre contraction, output-stationary.
//
// C_q[row, jj] = sum over degrees l of parity class q of A[row, l] P[l, jj]
// (same contraction as legendre.hip): the northern node gets C_E + C_O, its
// southern mirror C_E - C_O.
//
// Like the forward kernel (legendre_rows.hip) this is HBM-bound, so the
// layout reads the modal rows once and writes the Fourier rows once:
//   * a workgroup owns 64 rows (4 waves x 16) of one zonal wavenumber m and
//     keeps the accumulators of every node tile in VGPRs while it sweeps the
//     32-degree tiles of that m;
//   * per tile, each row of the workgroup brings in one whole 128-byte line
//     of its modal row (float4 loads; TL rows hold an odd number of degrees,
//     so rows start anywhere in a line) into a two-line LDS ring per row,
//     from which the waves read their A-operand fragments;
//   * the Legendre values arrive through a double-buffered LDS tile filled
//     from a plan table stored in fragment order (inverse_fragments_kernel);
//   * at the end the up / down halves are staged in LDS and written as whole
//     contiguous row segments.
//
// Reduction order inside a 32-degree tile lt .. lt + 31: lane group g =
// lane / 16 holds the degrees lt + 8 g .. lt + 8 g + 7; class q (d-parity)
// takes the degrees lt + 8 g + 2 e + q.  FP32 uses v_mfma_f32_16x16x4f32
// (e = u, one MFMA per u), the split path v_mfma_f32_16x16x16bf16_1k (e =
// 0..3 in one MFMA, three products hi*hi + hi*lo + lo*hi).
#include "dino_common.h"

namespace dino {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short shortx4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 64;
constexpr int kTile = 32;
constexpr int kRingPitch = 68;  // modal ring row pitch (floats): 64 degrees + 4

__device__ __forceinline__ void row_field(int m, int q, int* b, int* r) {
  if (m == 0) {
    *b = q;
    *r = 0;
  } else {
    *b = q >> 1;
    *r = 2 * m - 1 + (q & 1);
  }
}

static_assert(kTile == kLegTile, "tile width shared with the modal kernels");
__host__ __device__ inline int inv_tiles_of_m(int m, int L) { return leg_tile_count(m, L); }

// float4 units per degree tile: NT node tiles x 2 classes x 64 lanes.
__host__ __device__ constexpr int inv_tile_f4(int nt) { return 128 * nt; }

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  return __builtin_bit_cast(unsigned short, (__bf16)x);
}

__device__ __forceinline__ float bf16_value(unsigned short h) {
  return (float)__builtin_bit_cast(__bf16, h);
}

// VOUT: J % 8 == 0 and a 16-byte aligned Fourier buffer (float4 stores).
// WS: waves per 16-row strip (1 or 2).  With 2 the two waves of a strip split
// the node tiles, the modal loads and the output stores between them, so a
// block has 8 waves that each hold half the accumulators: twice the waves
// per CU to overlap the MFMAs with the loads and stores of other waves (the
// FP32 register budget of 6 waves per SIMD lets 3 such blocks share a CU;
// the split path would spill under it).
template <bool BF, int NT, bool VOUT, int WS>
__global__ __launch_bounds__(256 * WS, WS == 2 && !BF ? 6 : 1) void leg_inv_rows_kernel(
    const float* __restrict__ modal, float* __restrict__ fourier,
    const float4* __restrict__ itab, const int64_t* __restrict__ itile_off, int B, int L,
    int J, int Jh, int Rrows, int mpair) {
  static_assert(NT % WS == 0, "the waves of a strip split the node tiles evenly");
  constexpr int NTH = 256 * WS;  // threads
  constexpr int NTW = NT / WS;   // node tiles per wave
  constexpr int RPL = 2 / WS;    // modal line quads per lane per tile
  constexpr int TF4 = inv_tile_f4(NT);
  constexpr int PER = (TF4 + NTH - 1) / NTH;
  constexpr int OPITCH = 16 * NT + 4;
  constexpr int OUT4 = 4 * 16 * OPITCH / 4;
  constexpr int MAIN4 = (2 * TF4 > OUT4) ? 2 * TF4 : OUT4;
  __shared__ floatx4 lds[MAIN4];
  // Modal rows arrive as whole 128-byte lines (float4 loads, one new line
  // per row and tile) into a per-row ring of two lines: degree l of a row
  // sits at slot (ra + l) & 63, ra = row address / 4 mod 64, so a line is
  // contiguous there, and the tile lt .. lt + 31 reads the two lines that
  // hold it.  A TL row holds an odd number of degrees, so tile-sized
  // segments would start anywhere in a line and split into partial
  // requests.
  __shared__ floatx4 ring4[4 * 16 * kRingPitch / 4];
  float* const ring = reinterpret_cast<float*>(ring4);
  const int m = m_of_block(blockIdx.y, gridDim.y, mpair);
  const int nrows = (m == 0) ? B : 2 * B;
  const int row0 = blockIdx.x * kRows;
  if (row0 >= nrows) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int strip = wave & 3, wh = wave >> 2;  // wh: this wave's share of the strip
  const int g = lane >> 4, c16 = lane & 15;
  const int lt0 = leg_tile_start(m, L);
  const int ntile = inv_tiles_of_m(m, L);
  const float4* tb = itab + itile_off[m];
  float* const rs = ring + strip * 16 * kRingPitch;  // this strip's rings

  // Line loads of this lane: quad q of row rl of the strip (item = 8 rl + q;
  // the strip's 128 items split over its WS waves).  Line k of a row starts
  // at degree lt0 - a + 32 k, a = (ra + lt0) & 31: line 0 holds lt0.  Rows
  // past the batch load the last row, and lines past the row's last degree
  // the row's last line: every load is a whole line that holds an element
  // of the modal array (no address outside its pages), and what such a line
  // brings in is zeroed where it is read.
  const float* const fbase = modal + (int64_t)(m == 0 ? row0 : row0 >> 1) * Rrows * L;
  int loff[RPL], lmax[RPL], lslot[RPL];
#pragma unroll
  for (int i = 0; i < RPL; ++i) {
    const int item = WS == 2 ? 64 * wh + lane : lane + 64 * i;
    const int rl = item >> 3, q = item & 7;
    const int row = row0 + strip * 16 + rl;
    int b, r;
    row_field(m, row < nrows ? row : nrows - 1, &b, &r);
    const int off = ((b - (m == 0 ? row0 : row0 >> 1)) * Rrows + r) * L;  // < 2^31 (checked)
    const int ra = (int)((reinterpret_cast<uintptr_t>(fbase + off) >> 2) & 63);
    const int a = (ra + lt0) & 31;
    loff[i] = off + lt0 - a + 4 * q;         // quad q of line 0
    lmax[i] = (L - 1 - (lt0 - a)) >> 5;      // last line holding a degree < L
    lslot[i] = (((ra + lt0) & 32) + 4 * q) | (rl << 8);  // ring slot of quad q, line 0; row
  }
  auto load_line = [&](floatx4* dst, int k) {
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int kk = k < lmax[i] ? k : lmax[i];
      dst[i] = *reinterpret_cast<const floatx4*>(fbase + loff[i] + 32 * kk);
    }
  };
  auto store_line = [&](const floatx4* src, int k) {
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int rl = lslot[i] >> 8;
      *reinterpret_cast<floatx4*>(rs + rl * kRingPitch + (((lslot[i] & 255) + 32 * k) & 63)) =
          src[i];
    }
  };
  // reader side: row c16 of the strip
  int ra_c;
  bool rok_c;
  {
    const int row = row0 + strip * 16 + c16;
    rok_c = row < nrows;
    int b, r;
    row_field(m, rok_c ? row : nrows - 1, &b, &r);
    const int off = ((b - (m == 0 ? row0 : row0 >> 1)) * Rrows + r) * L;
    ra_c = (int)((reinterpret_cast<uintptr_t>(fbase + off) >> 2) & 63);
  }

  // prefetch registers: table tile, and the strip's next modal line
  floatx4 st[PER], ln0[RPL], ln1[RPL];
  auto load_table = [&](int tt) {
    const float4* src = tb + (int64_t)tt * TF4;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = tid + NTH * i;
      st[i] = *reinterpret_cast<const floatx4*>(src + (idx < TF4 ? idx : 0));
    }
  };
  load_table(0);
  load_line(ln0, 0);
  load_line(ln1, 1);

  floatx4 acc[NTW][2];
#pragma unroll
  for (int n = 0; n < NTW; ++n) acc[n][0] = acc[n][1] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int tt = 0; tt < ntile; ++tt) {
    floatx4* buf = lds + (tt & 1) * TF4;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = tid + NTH * i;
      if (idx < TF4) buf[idx] = st[i];
    }
    // the ring holds lines tt and tt + 1 from here on
    if (tt == 0) store_line(ln0, 0);
    store_line(ln1, tt + 1);
    __syncthreads();
    // the last iteration reloads its own tile, which is never stored
    load_table(tt + 1 < ntile ? tt + 1 : tt);
    load_line(ln1, tt + 2);
    // keep the prefetch ahead of the MFMAs (the scheduler would otherwise
    // sink the loads to the end of the tile, serialising them with it)
    __builtin_amdgcn_sched_barrier(0);
    const int lt = lt0 + kTile * tt;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int l = lt + 8 * g + e;
      const float x = rs[c16 * kRingPitch + ((ra_c + l) & 63)];
      v[e] = (rok_c && l < L) ? x : 0.f;
    }
    // both waves of a strip have read line tt before either writes line
    // tt + 2 over it (next iteration)
    if constexpr (WS == 2) __syncthreads();
    else wave_lds_sync();
    if constexpr (BF) {
      shortx4 hi[2], lo[2];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float x = v[2 * e + q];
          const unsigned short h = bf16_bits(x);
          hi[q][e] = (short)h;
          lo[q][e] = (short)bf16_bits(x - bf16_value(h));
        }
#pragma unroll
      for (int n = 0; n < NTW; ++n)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const floatx4 p = buf[((wh * NTW + n) * 2 + q) * 64 + lane];
          const shortx4 ph = __builtin_bit_cast(shortx4, __builtin_shufflevector(p, p, 0, 1));
          const shortx4 pl = __builtin_bit_cast(shortx4, __builtin_shufflevector(p, p, 2, 3));
          acc[n][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lo[q], ph, acc[n][q], 0, 0, 0);
          acc[n][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(hi[q], pl, acc[n][q], 0, 0, 0);
          acc[n][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(hi[q], ph, acc[n][q], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int n = 0; n < NTW; ++n)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const floatx4 p = buf[((wh * NTW + n) * 2 + q) * 64 + lane];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            acc[n][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[2 * u + q], p[u], acc[n][q], 0, 0, 0);
        }
    }
  }

  // Epilogue: class E (even l - m) is the d-parity q = (lt0 - m) & 1.
  const int qe = (lt0 - m) & 1;
  __syncthreads();  // every wave is done with the table buffers
  float* os = reinterpret_cast<float*>(lds) + strip * 16 * OPITCH;
  const int ndn = J - Jh;  // southern nodes (the equator of odd J is northern)
  // the output staging rows belong to one strip: its own wave, or two waves
  auto strip_sync = [&]() {
    if constexpr (WS == 1) wave_lds_sync();
    else __syncthreads();
  };
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int len = half == 0 ? Jh : ndn;
#pragma unroll
    for (int n = 0; n < NTW; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int jj = 16 * (wh * NTW + n) + c16;
        // selects, not a runtime array index (which would demote acc to scratch)
        const float a0 = acc[n][0][i], a1 = acc[n][1][i];
        const float e = qe ? a1 : a0, o = qe ? a0 : a1;
        if (half == 0) {
          if (jj < Jh) os[(4 * g + i) * OPITCH + jj] = e + o;
        } else {
          const int pos = Jh - 1 - jj;  // position in the southern segment
          if (jj < Jh && pos < ndn) os[(4 * g + i) * OPITCH + pos] = e - o;
        }
      }
    strip_sync();
    const int col0 = half == 0 ? J - Jh : 0;
    if constexpr (VOUT) {
      const int n4 = len / 4;  // J % 8 == 0: both halves hold Jh floats
      for (int idx = lane + 64 * wh; idx < 16 * n4; idx += 64 * WS) {
        const int rl = idx / n4, c = idx - rl * n4;
        const int row = row0 + strip * 16 + rl;
        if (row < nrows) {
          int b, r;
          row_field(m, row, &b, &r);
          *reinterpret_cast<floatx4*>(fourier + ((int64_t)b * Rrows + r) * J + col0 + 4 * c) =
              *reinterpret_cast<const floatx4*>(os + rl * OPITCH + 4 * c);
        }
      }
    } else {
      for (int idx = lane + 64 * wh; idx < 16 * len; idx += 64 * WS) {
        const int rl = idx / len, c = idx - rl * len;
        const int row = row0 + strip * 16 + rl;
        if (row < nrows) {
          int b, r;
          row_field(m, row, &b, &r);
          fourier[((int64_t)b * Rrows + r) * J + col0 + c] = os[rl * OPITCH + c];
        }
      }
    }
    strip_sync();
  }
}

// One thread per (tile, node tile, class, lane) slot of one m (blockIdx.y).
__global__ void inverse_fragments_kernel(int L, int Jh, int Jhp, int NT,
                                         const double* __restrict__ p64,
                                         const int64_t* __restrict__ m_off,
                                         const int64_t* __restrict__ itile_off,
                                         float* __restrict__ f32, float* __restrict__ fbf) {
  const int m = blockIdx.y;
  const int64_t per_tile = inv_tile_f4(NT);  // slots = float4 per tile
  const int64_t n = per_tile * inv_tiles_of_m(m, L);
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int tile = (int)(idx / per_tile);
  const int rem = (int)(idx - tile * per_tile);
  const int nt = rem / 128, q = (rem / 64) % 2, lane = rem % 64;
  const int lt = leg_tile_start(m, L) + kTile * tile;
  const int gg = lane / 16, jj = 16 * nt + lane % 16;
  const int64_t out = (itile_off[m] + idx) * 4;  // floats
  unsigned short h[4], lo[4];
  for (int e = 0; e < 4; ++e) {
    const int l = lt + 8 * gg + 2 * e + q;
    const double v =
        (l >= m && l < L && jj < Jh) ? p64[m_off[m] + (int64_t)(l - m) * Jhp + jj] : 0.0;
    f32[out + e] = (float)v;
    const __bf16 bh = (__bf16)(float)v;
    h[e] = __builtin_bit_cast(unsigned short, bh);
    lo[e] = __builtin_bit_cast(unsigned short, (__bf16)(float)(v - (double)(float)bh));
  }
  // BF16 slot: hi[0..3] then lo[0..3] (16 bytes)
  uint32_t* w = reinterpret_cast<uint32_t*>(fbf + out);
  w[0] = h[0] | ((uint32_t)h[1] << 16);
  w[1] = h[2] | ((uint32_t)h[3] << 16);
  w[2] = lo[0] | ((uint32_t)lo[1] << 16);
  w[3] = lo[2] | ((uint32_t)lo[3] << 16);
}

template <bool BF, int NT>
int launch_inv_rows(const dino_plan* pl, const float* in, float* out, int64_t B, hipStream_t s) {
  const bool vout = pl->J % 8 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0;
  const dim3 grid((unsigned)((2 * B + kRows - 1) / kRows), (unsigned)pl->M);
  const float4* tab = BF ? reinterpret_cast<const float4*>(pl->ifragbf)
                         : reinterpret_cast<const float4*>(pl->ifrag32);
  const int mpair = (pl->sched[DINO_STAGE_LEGENDRE_INVERSE] & DINO_SCHED_MPAIRED) ? 1 : 0;
  // two waves per strip whenever the node tiles split evenly
  constexpr int WS = NT % 2 == 0 ? 2 : 1;
  if (vout)
    leg_inv_rows_kernel<BF, NT, true, WS><<<grid, 256 * WS, 0, s>>>(
        in, out, tab, pl->itile_off, (int)B, pl->L, pl->J, pl->Jh, pl->R, mpair);
  else
    leg_inv_rows_kernel<BF, NT, false, WS><<<grid, 256 * WS, 0, s>>>(
        in, out, tab, pl->itile_off, (int)B, pl->L, pl->J, pl->Jh, pl->R, mpair);
  DINO_LAUNCH_CHECK("legendre inverse rows kernel");
  return DINO_OK;
}

template <bool BF>
int inv_rows_by_nt(const dino_plan* pl, const float* in, float* out, int64_t B, hipStream_t s) {
  switch (pl->int_nt) {
    case 1: return launch_inv_rows<BF, 1>(pl, in, out, B, s);
    case 2: return launch_inv_rows<BF, 2>(pl, in, out, B, s);
    case 4: return launch_inv_rows<BF, 4>(pl, in, out, B, s);
    case 8: return launch_inv_rows<BF, 8>(pl, in, out, B, s);
    default:
      set_error("inverse fragment tables hold %d node tiles", pl->int_nt);
      return DINO_EINVAL;
  }
}

}  // namespace

// Node tiles of 16 the inverse tables use (1, 2, 4 or 8; 0 when Jh > 128 and
// the tiled kernels are used instead).
int inverse_fragment_tiles(int Jh) {
  int nt = 1;
  while (16 * nt < Jh) nt *= 2;
  return nt <= 8 ? nt : 0;
}

int64_t inverse_fragment_layout(int M, int L, int nt, int64_t* off) {
  int64_t acc = 0;
  for (int m = 0; m < M; ++m) {
    off[m] = acc;
    acc += (int64_t)inv_tiles_of_m(m, L) * inv_tile_f4(nt);
  }
  return acc;
}

void inverse_fragment_fill(const dino_plan* pl, hipStream_t s) {
  const int64_t most = (int64_t)inv_tile_f4(pl->int_nt) * inv_tiles_of_m(0, pl->L);
  const dim3 grid((unsigned)((most + 255) / 256), (unsigned)pl->M);
  inverse_fragments_kernel<<<grid, 256, 0, s>>>(pl->L, pl->Jh, pl->Jhp, pl->int_nt, pl->p64,
                                                pl->m_offset, pl->itile_off, pl->ifrag32,
                                                pl->ifragbf);
}

int legendre_inverse_rows(const dino_plan* pl, const float* in, float* out, int64_t B,
                          bool split, hipStream_t s) {
  if (B == 0) return DINO_OK;
  DINO_REQUIRE(2 * B / kRows < (int64_t)1 << 30, DINO_EINVAL, "batch too large");
  // a block's modal rows span at most 64 fields, addressed with 32-bit offsets
  DINO_REQUIRE((int64_t)kRows * pl->R * pl->L < (int64_t)1 << 31, DINO_EINVAL,
               "modal fields too large for the row kernel");
  return split ? inv_rows_by_nt<true>(pl, in, out, B, s) : inv_rows_by_nt<false>(pl, in, out, B, s);
}

}  // namespace dino
