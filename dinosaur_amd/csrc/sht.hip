// C-ABI entry points of the spherical harmonic transform: the longitude DFT
// (dft.hip) and the folded Legendre GEMMs (legendre*.hip) chained through a
// caller-provided workspace holding the (batch, 2M-1, J) Fourier rows.
#include "dino_common.h"

#include <mutex>

// Fourier rows of dino_sht_inverse_scratch.  Every GPU holds its own copy of
// this array until the process ends.  In host code the name is a stand-in,
// not a device address: host code uses only the per-device address that
// scratch_address() obtains with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[DINO_SHT_SCRATCH_BYTES];

namespace dino {
int dft_forward_impl(const dino_plan*, const void*, void*, int64_t, int, int, hipStream_t);
int dft_inverse_impl(const dino_plan*, const void*, void*, int64_t, int, int, hipStream_t);
int legendre_inverse_fma(const dino_plan*, const void*, void*, int64_t, int, hipStream_t);
int legendre_forward_fma(const dino_plan*, const void*, void*, int64_t, int, hipStream_t);
int legendre_inverse_mfma(const dino_plan*, const void*, void*, int64_t, int, hipStream_t);
int legendre_forward_mfma(const dino_plan*, const void*, void*, int64_t, int, bool, hipStream_t);
}  // namespace dino

using namespace dino;

namespace {

int check_common(const dino_plan* p, const void* a, const void* b, int64_t batch, int dtype,
                 int precision) {
  DINO_REQUIRE(p != nullptr, DINO_EINVAL, "NULL plan");
  DINO_REQUIRE(batch >= 0, DINO_EINVAL, "negative batch");
  DINO_REQUIRE(batch == 0 || (a && b), DINO_EINVAL, "NULL buffer");
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "unknown dtype %d", dtype);
  DINO_REQUIRE(precision == DINO_PREC_NATIVE || precision == DINO_PREC_BF16X3,
               DINO_EINVAL, "unknown precision %d", precision);
  DINO_REQUIRE(precision == DINO_PREC_NATIVE || dtype == DINO_F32, DINO_EINVAL,
               "split-precision transforms take FP32 data");
  return DINO_OK;
}

// FP64 data: FP64 FMA tiles.  FP32 data: MFMA (FP32 or split BF16 x3).
int leg_inverse(const dino_plan* p, const void* in, void* out, int64_t B, int dtype, int prec,
                hipStream_t s) {
  if (dtype == DINO_F64) return legendre_inverse_fma(p, in, out, B, dtype, s);
  return legendre_inverse_mfma(p, in, out, B, prec, s);
}

// triangle_only: rows may skip the all-zero degrees below leg_tile_start(m, L)
// (the FP64 and tiled kernels write them anyway, which is also allowed).
int leg_forward(const dino_plan* p, const void* in, void* out, int64_t B, int dtype, int prec,
                hipStream_t s, bool triangle_only = false) {
  if (dtype == DINO_F64) return legendre_forward_fma(p, in, out, B, dtype, s);
  return legendre_forward_mfma(p, in, out, B, prec, triangle_only, s);
}

// Splits the DINO_SHT_TRIANGLE_ONLY flag off a precision argument.
bool take_triangle_flag(int* precision) {
  const bool t = (*precision & DINO_SHT_TRIANGLE_ONLY) != 0;
  *precision &= ~DINO_SHT_TRIANGLE_ONLY;
  return t;
}

// Device address of `scratch` on the current device, looked up once per
// device (each device has its own copy at its own address).
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices];

int scratch_address(void** out) {
  int dev = -1;
  DINO_HIP(hipGetDevice(&dev));
  DINO_REQUIRE(dev >= 0 && dev < kScratchDevices, DINO_EINVAL,
               "device %d is beyond the %d devices the scratch table holds", dev,
               kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    DINO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DINO_REQUIRE(p != nullptr, DINO_EHIP, "hipGetSymbolAddress(scratch) gave NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DINO_OK;
}

}  // namespace

#define CHECK_OR_RETURN(expr)    \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != DINO_OK) return rc_; \
  } while (0)

extern "C" int dino_sht_inverse(const dino_plan* p, const void* modal, void* nodal, void* work,
                                int64_t batch, int dtype, int precision, void* stream) {
  CHECK_OR_RETURN(check_common(p, modal, nodal, batch, dtype, precision));
  DINO_REQUIRE(batch == 0 || work, DINO_EINVAL, "NULL workspace");
  hipStream_t s = as_stream(stream);
  CHECK_OR_RETURN(leg_inverse(p, modal, work, batch, dtype, precision, s));
  return dft_inverse_impl(p, work, nodal, batch, dtype, 0, s);
}

// The same kernels as dino_sht_inverse, with the Fourier rows in `scratch`.
// The Legendre stage writes every (field, row, node) entry of the rows before
// the DFT stage reads them, so nothing needs clearing between calls.
extern "C" int dino_sht_inverse_scratch(const dino_plan* p, const void* modal, void* nodal,
                                        int64_t batch, int dtype, int precision, void* stream) {
  CHECK_OR_RETURN(check_common(p, modal, nodal, batch, dtype, precision));
  if (batch == 0) return DINO_OK;
  const size_t per_field = (dtype == DINO_F64 ? 8 : 4) * (size_t)p->R * (size_t)p->J;
  DINO_REQUIRE((uint64_t)batch <= DINO_SHT_SCRATCH_BYTES / per_field, DINO_EINVAL,
               "batch %lld needs %llu bytes of Fourier rows; the scratch array holds %llu",
               (long long)batch, (unsigned long long)per_field * (unsigned long long)batch,
               (unsigned long long)DINO_SHT_SCRATCH_BYTES);
  void* work = nullptr;
  CHECK_OR_RETURN(scratch_address(&work));
  hipStream_t s = as_stream(stream);
  CHECK_OR_RETURN(leg_inverse(p, modal, work, batch, dtype, precision, s));
  return dft_inverse_impl(p, work, nodal, batch, dtype, 0, s);
}

extern "C" int dino_sht_forward(const dino_plan* p, const void* nodal, void* modal, void* work,
                                int64_t batch, int dtype, int precision, void* stream) {
  const bool tri = take_triangle_flag(&precision);
  CHECK_OR_RETURN(check_common(p, nodal, modal, batch, dtype, precision));
  DINO_REQUIRE(batch == 0 || work, DINO_EINVAL, "NULL workspace");
  hipStream_t s = as_stream(stream);
  CHECK_OR_RETURN(dft_forward_impl(p, nodal, work, batch, dtype, 1, s));
  return leg_forward(p, work, modal, batch, dtype, precision, s, tri);
}

extern "C" int dino_sht_inverse_adjoint(const dino_plan* p, const void* nodal_cot,
                                        void* modal_cot, void* work, int64_t batch, int dtype,
                                        int precision, void* stream) {
  const bool tri = take_triangle_flag(&precision);
  CHECK_OR_RETURN(check_common(p, nodal_cot, modal_cot, batch, dtype, precision));
  DINO_REQUIRE(batch == 0 || work, DINO_EINVAL, "NULL workspace");
  hipStream_t s = as_stream(stream);
  CHECK_OR_RETURN(dft_forward_impl(p, nodal_cot, work, batch, dtype, 0, s));
  return leg_forward(p, work, modal_cot, batch, dtype, precision, s, tri);
}

extern "C" int dino_sht_forward_adjoint(const dino_plan* p, const void* modal_cot,
                                        void* nodal_cot, void* work, int64_t batch, int dtype,
                                        int precision, void* stream) {
  CHECK_OR_RETURN(check_common(p, modal_cot, nodal_cot, batch, dtype, precision));
  DINO_REQUIRE(batch == 0 || work, DINO_EINVAL, "NULL workspace");
  hipStream_t s = as_stream(stream);
  CHECK_OR_RETURN(leg_inverse(p, modal_cot, work, batch, dtype, precision, s));
  return dft_inverse_impl(p, work, nodal_cot, batch, dtype, 1, s);
}

extern "C" int dino_dft_forward(const dino_plan* p, const void* nodal, void* fourier,
                                int64_t batch, int dtype, int weighted, void* stream) {
  CHECK_OR_RETURN(check_common(p, nodal, fourier, batch, dtype, DINO_PREC_NATIVE));
  return dft_forward_impl(p, nodal, fourier, batch, dtype, weighted, as_stream(stream));
}

extern "C" int dino_dft_inverse(const dino_plan* p, const void* fourier, void* nodal,
                                int64_t batch, int dtype, int weighted, void* stream) {
  CHECK_OR_RETURN(check_common(p, fourier, nodal, batch, dtype, DINO_PREC_NATIVE));
  return dft_inverse_impl(p, fourier, nodal, batch, dtype, weighted, as_stream(stream));
}

extern "C" int dino_legendre_inverse(const dino_plan* p, const void* modal, void* fourier,
                                     int64_t batch, int dtype, int precision, void* stream) {
  CHECK_OR_RETURN(check_common(p, modal, fourier, batch, dtype, precision));
  return leg_inverse(p, modal, fourier, batch, dtype, precision, as_stream(stream));
}

extern "C" int dino_legendre_forward(const dino_plan* p, const void* fourier, void* modal,
                                     int64_t batch, int dtype, int precision, void* stream) {
  const bool tri = take_triangle_flag(&precision);
  CHECK_OR_RETURN(check_common(p, fourier, modal, batch, dtype, precision));
  return leg_forward(p, fourier, modal, batch, dtype, precision, as_stream(stream), tri);
}
