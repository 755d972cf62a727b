// Plan creation: FP64 Gauss nodes and the folded, normalised P_l^m table are
// generated on the GPU; the FP32 copy and the longitude twiddles are derived
// from them on the GPU.
//
// Table layout ("folded triangular"): for each m, rows l = m .. L-1, each row
// holds Jh = ceil(J/2) values P_l^m(x_j) for the northern nodes
// j = J - Jh .. J - 1, padded to Jhp.  The southern values follow from
// P_l^m(-x) = (-1)^(l+m) P_l^m(x), so the table is a quarter of the
// reference's dense (2M-1, J, L) array (spherical_harmonic.py:250-259).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "dino_common.h"

namespace dino {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int hip_fail(hipError_t e, const char* what) {
  set_error("HIP error %d (%s) in %s", static_cast<int>(e), hipGetErrorString(e),
            what);
  return e == hipErrorOutOfMemory ? DINO_ENOMEM : DINO_EHIP;
}

// k-th Gauss-Legendre root in descending order, by Newton's method on the
// three-term recurrence P_{j+1} = ((2j+1) x P_j - j P_{j-1}) / (j+1).
__host__ __device__ inline void gauss_root(int n, int k, double* x_out,
                                           double* w_out) {
  if (n % 2 == 1 && k == n / 2) {
    // P_n is odd for odd n: the middle root is exactly zero.
    double p0 = 1.0, p1 = 0.0;
    for (int j = 1; j < n; ++j) {
      double p2 = ((2.0 * j + 1.0) * 0.0 * p1 - j * p0) / (j + 1.0);
      p0 = p1;
      p1 = p2;
    }
    double dp = n * (0.0 * p1 - p0) / (0.0 - 1.0);
    *x_out = 0.0;
    *w_out = 2.0 / (dp * dp);
    return;
  }
  const double pi = 3.14159265358979323846;
  double x = cos(pi * (k + 0.75) / (n + 0.5));
  int polish = 0;
  for (int it = 0; it < 100; ++it) {
    double p0 = 1.0, p1 = x;
    for (int j = 1; j < n; ++j) {
      double p2 = ((2.0 * j + 1.0) * x * p1 - j * p0) / (j + 1.0);
      p0 = p1;
      p1 = p2;
    }
    double dp = n * (x * p1 - p0) / (x * x - 1.0);
    double dx = p1 / dp;
    x -= dx;
    if (fabs(dx) <= 1e-15) {
      if (++polish >= 2) break;
    }
  }
  double p0 = 1.0, p1 = x;
  for (int j = 1; j < n; ++j) {
    double p2 = ((2.0 * j + 1.0) * x * p1 - j * p0) / (j + 1.0);
    p0 = p1;
    p1 = p2;
  }
  double dp = n * (x * p1 - p0) / (x * x - 1.0);
  *x_out = x;
  *w_out = 2.0 / ((1.0 - x * x) * dp * dp);
}

__global__ void gauss_nodes_kernel(int n, double* x, double* w) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int half = (n + 1) / 2;
  if (k >= half) return;
  double xr, wr;
  gauss_root(n, k, &xr, &wr);
  // ascending order with exact mirror symmetry
  x[n - 1 - k] = xr;
  w[n - 1 - k] = wr;
  x[k] = -xr;
  w[k] = wr;
}

// One thread per (m, jj): sectoral start then the three-term recurrence in
// k = l - m, in the operation order of associated_legendre.py:51-67 (no FMA
// contraction, so a table built from identical nodes is bitwise identical).
#pragma clang fp contract(off)
__global__ void legendre_table_kernel(int M, int L, int J, int Jh, int Jhp,
                                      const double* __restrict__ nodes,
                                      const int64_t* __restrict__ m_off,
                                      double* __restrict__ p) {
  int jj = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (jj >= Jhp || m >= M) return;
  double* col = p + m_off[m] + jj;
  if (jj >= Jh) {
    for (int l = m; l < L; ++l) col[(int64_t)(l - m) * Jhp] = 0.0;
    return;
  }
  double x = nodes[J - Jh + jj];
  double y = sqrt(1.0 - x * x);
  double sect = 1.0 / sqrt(2.0);
  for (int k = 1; k <= m; ++k) sect = (-sqrt(1.0 + 1.0 / (2.0 * k)) * y) * sect;
  double pkm2 = 0.0, pkm1 = sect;
  col[0] = sect;
  const int64_t m2 = (int64_t)m * m;
  for (int l = m + 1; l < L; ++l) {
    int64_t mk2 = (int64_t)l * l;
    int64_t mkp2 = (int64_t)(l - 1) * (l - 1);
    double a = sqrt((double)(4 * mk2 - 1) / (double)(mk2 - m2));
    double b = sqrt((double)(mkp2 - m2) / (double)(4 * mkp2 - 1));
    double cur = a * (x * pkm1 - b * pkm2);
    col[(int64_t)(l - m) * Jhp] = cur;
    pkm2 = pkm1;
    pkm1 = cur;
  }
}
#pragma clang fp contract(on)

__global__ void round_table_kernel(int64_t n, const double* __restrict__ p64,
                                   float* __restrict__ p32) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p32[i] = (float)p64[i];
}

__global__ void twiddle_kernel(int n, double* tw64, float* tw32) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(2.0 * k / n, &s, &c);
  tw64[2 * k] = c;
  tw64[2 * k + 1] = s;
  tw32[2 * k] = (float)c;
  tw32[2 * k + 1] = (float)s;
}

__global__ void quad_weights_kernel(int I, int J, const double* w, double* q64,
                                    float* q32) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  const double pi = 3.14159265358979323846;
  double q = (2.0 * pi / I) * w[j];
  q64[j] = q;
  q32[j] = (float)q;
}

int forward_fragment_steps(int Jh);
int64_t forward_fragment_layout(int M, int L, int ks, int64_t* off);
void forward_fragment_fill(const dino_plan* pl, hipStream_t s);
int inverse_fragment_tiles(int Jh);
int64_t inverse_fragment_layout(int M, int L, int nt, int64_t* off);
void inverse_fragment_fill(const dino_plan* pl, hipStream_t s);

static void free_plan(dino_plan* p) {
  if (!p) return;
  void* bufs[] = {p->nodes64, p->weights64, p->qw32, p->qw64, p->m_offset,
                  p->p64, p->p32,
                  p->tw64, p->tw32, p->ftile_off, p->ffrag32, p->ffragbf,
                  p->itile_off, p->ifrag32, p->ifragbf, p->sink};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete p;
}

}  // namespace dino

using namespace dino;

extern "C" const char* dino_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* dino_version(void) { return "dinosaur_amd 0.2 gfx950 mfma-f32 split-bf16x3"; }

extern "C" int dino_gauss_nodes_host(int n, double* nodes, double* weights) {
  DINO_REQUIRE(n >= 1 && nodes && weights, DINO_EINVAL,
               "dino_gauss_nodes_host: invalid arguments (n=%d)", n);
  for (int k = 0; k < (n + 1) / 2; ++k) {
    double x, w;
    gauss_root(n, k, &x, &w);
    nodes[n - 1 - k] = x;
    weights[n - 1 - k] = w;
    nodes[k] = -x;
    weights[k] = w;
  }
  return DINO_OK;
}

extern "C" int dino_plan_create(int M, int L, int I, int J, const double* nodes,
                                const double* weights, dino_plan** out) {
  DINO_REQUIRE(out != nullptr, DINO_EINVAL, "dino_plan_create: out is NULL");
  *out = nullptr;
  DINO_REQUIRE(M >= 1 && L >= M, DINO_EINVAL,
               "Expected n_m <= n_l; got n_m = %d and n_l = %d.", M, L);
  DINO_REQUIRE(I >= 2 * M - 1 && I >= M, DINO_EINVAL,
               "real basis requires longitude nodes >= 2*wavenumbers-1 for an "
               "exact real transform; got I = %d, M = %d", I, M);
  DINO_REQUIRE(J >= 1, DINO_EINVAL, "latitude_nodes must be positive");
  DINO_REQUIRE((nodes == nullptr) == (weights == nullptr), DINO_EINVAL,
               "nodes and weights must both be given or both be NULL");
  if (nodes) {
    for (int j = 0; j < J; ++j)
      DINO_REQUIRE(nodes[j] == -nodes[J - 1 - j], DINO_EINVAL,
                   "latitude nodes must be exactly symmetric about 0");
  }
  dino_plan* p = new dino_plan();
  std::memset(p, 0, sizeof(*p));
  p->M = M;
  p->L = L;
  p->I = I;
  p->J = J;
  p->R = 2 * M - 1;
  p->Jh = (J + 1) / 2;
  p->Jhp = (p->Jh + 7) / 8 * 8;
  p->pow2 = (I & (I - 1)) == 0 && I >= 8 && I <= 4096;
  std::vector<int64_t> off(M);
  int64_t acc = 0;
  for (int m = 0; m < M; ++m) {
    off[m] = acc;
    acc += (int64_t)(L - m) * p->Jhp;
  }
  p->table_elems = acc;
  auto fail = [&](hipError_t e, const char* what) {
    int rc = hip_fail(e, what);
    free_plan(p);
    return rc;
  };
#define PLAN_HIP(call)                      \
  do {                                      \
    hipError_t e_ = (call);                 \
    if (e_ != hipSuccess) return fail(e_, #call); \
  } while (0)
  PLAN_HIP(hipMalloc(&p->nodes64, sizeof(double) * J));
  PLAN_HIP(hipMalloc(&p->weights64, sizeof(double) * J));
  PLAN_HIP(hipMalloc(&p->qw64, sizeof(double) * J));
  PLAN_HIP(hipMalloc(&p->qw32, sizeof(float) * J));
  PLAN_HIP(hipMalloc(&p->m_offset, sizeof(int64_t) * M));
  PLAN_HIP(hipMalloc(&p->p64, sizeof(double) * acc));
  PLAN_HIP(hipMalloc(&p->p32, sizeof(float) * acc));
  PLAN_HIP(hipMalloc(&p->tw64, sizeof(double) * 2 * I));
  PLAN_HIP(hipMalloc(&p->tw32, sizeof(float) * 2 * I));
  PLAN_HIP(hipMemcpy(p->m_offset, off.data(), sizeof(int64_t) * M,
                     hipMemcpyHostToDevice));
  if (nodes) {
    PLAN_HIP(hipMemcpy(p->nodes64, nodes, sizeof(double) * J, hipMemcpyHostToDevice));
    PLAN_HIP(hipMemcpy(p->weights64, weights, sizeof(double) * J, hipMemcpyHostToDevice));
  } else {
    int half = (J + 1) / 2;
    gauss_nodes_kernel<<<(half + 63) / 64, 64>>>(J, p->nodes64, p->weights64);
    PLAN_HIP(hipGetLastError());
  }
  quad_weights_kernel<<<(J + 255) / 256, 256>>>(I, J, p->weights64, p->qw64, p->qw32);
  PLAN_HIP(hipGetLastError());
  dim3 tgrid((p->Jhp + 63) / 64, M);
  legendre_table_kernel<<<tgrid, 64>>>(M, L, J, p->Jh, p->Jhp, p->nodes64,
                                        p->m_offset, p->p64);
  PLAN_HIP(hipGetLastError());
  round_table_kernel<<<(unsigned)((acc + 255) / 256), 256>>>(acc, p->p64, p->p32);
  PLAN_HIP(hipGetLastError());
  twiddle_kernel<<<(I + 255) / 256, 256>>>(I, p->tw64, p->tw32);
  PLAN_HIP(hipGetLastError());
  p->fks = forward_fragment_steps(p->Jh);
  if (p->fks > 0) {
    std::vector<int64_t> foff(M);
    const int64_t f4 = forward_fragment_layout(M, L, p->fks, foff.data());
    PLAN_HIP(hipMalloc(&p->ftile_off, sizeof(int64_t) * M));
    PLAN_HIP(hipMalloc(&p->ffrag32, 16 * f4));
    PLAN_HIP(hipMalloc(&p->ffragbf, 16 * f4));
    PLAN_HIP(hipMalloc(&p->sink, sizeof(float) * 256));
    PLAN_HIP(hipMemcpy(p->ftile_off, foff.data(), sizeof(int64_t) * M, hipMemcpyHostToDevice));
    forward_fragment_fill(p, 0);
    PLAN_HIP(hipGetLastError());
  }
  p->int_nt = inverse_fragment_tiles(p->Jh);
  if (p->int_nt > 0) {
    std::vector<int64_t> ioff(M);
    const int64_t f4 = inverse_fragment_layout(M, L, p->int_nt, ioff.data());
    PLAN_HIP(hipMalloc(&p->itile_off, sizeof(int64_t) * M));
    PLAN_HIP(hipMalloc(&p->ifrag32, 16 * f4));
    PLAN_HIP(hipMalloc(&p->ifragbf, 16 * f4));
    PLAN_HIP(hipMemcpy(p->itile_off, ioff.data(), sizeof(int64_t) * M, hipMemcpyHostToDevice));
    inverse_fragment_fill(p, 0);
    PLAN_HIP(hipGetLastError());
  }
  PLAN_HIP(hipDeviceSynchronize());
#undef PLAN_HIP
  *out = p;
  return DINO_OK;
}

extern "C" int dino_plan_destroy(dino_plan* plan) {
  free_plan(plan);
  return DINO_OK;
}

extern "C" int dino_plan_set_schedule(dino_plan* p, int stage, int schedule) {
  DINO_REQUIRE(p, DINO_EINVAL, "dino_plan_set_schedule: NULL plan");
  DINO_REQUIRE(stage >= DINO_STAGE_LEGENDRE_INVERSE && stage <= DINO_STAGE_DFT_FORWARD,
               DINO_EINVAL, "dino_plan_set_schedule: invalid stage %d", stage);
  const int fam = schedule & 0xff, flags = schedule & ~0xff;
  const bool legendre = stage <= DINO_STAGE_LEGENDRE_FORWARD;
  DINO_REQUIRE(schedule >= 0 && (flags & ~DINO_SCHED_MPAIRED) == 0 &&
                   (legendre || flags == 0),
               DINO_EINVAL, "dino_plan_set_schedule: invalid flags in schedule 0x%x", schedule);
  if (legendre) {
    DINO_REQUIRE(fam <= DINO_SCHED_ROWS_WIDE, DINO_EINVAL,
                 "dino_plan_set_schedule: invalid Legendre kernel family %d", fam);
    DINO_REQUIRE(fam != DINO_SCHED_ROWS_WIDE || stage == DINO_STAGE_LEGENDRE_FORWARD,
                 DINO_EINVAL, "dino_plan_set_schedule: the wide row kernel is forward only");
    if (fam == DINO_SCHED_ROWS || fam == DINO_SCHED_ROWS_WIDE) {
      const bool have = stage == DINO_STAGE_LEGENDRE_INVERSE ? p->int_nt > 0 : p->fks > 0;
      DINO_REQUIRE(have, DINO_EINVAL,
                   "dino_plan_set_schedule: no row-resident Legendre kernel for %d latitudes",
                   p->J);
    }
  } else {
    DINO_REQUIRE(fam <= DINO_SCHED_PER_TILE, DINO_EINVAL,
                 "dino_plan_set_schedule: invalid DFT kernel family %d", fam);
  }
  p->sched[stage] = schedule;
  return DINO_OK;
}

extern "C" int dino_plan_get_schedule(const dino_plan* p, int stage) {
  if (!p || stage < DINO_STAGE_LEGENDRE_INVERSE || stage > DINO_STAGE_DFT_FORWARD) return -1;
  return p->sched[stage];
}

extern "C" int dino_plan_nodes(const dino_plan* p, double* nodes, double* weights) {
  DINO_REQUIRE(p && nodes && weights, DINO_EINVAL, "dino_plan_nodes: NULL argument");
  DINO_HIP(hipMemcpy(nodes, p->nodes64, sizeof(double) * p->J, hipMemcpyDeviceToHost));
  DINO_HIP(hipMemcpy(weights, p->weights64, sizeof(double) * p->J, hipMemcpyDeviceToHost));
  return DINO_OK;
}

extern "C" int64_t dino_plan_table_size(const dino_plan* p) {
  if (!p) return -1;
  int64_t n = 0;
  for (int m = 0; m < p->M; ++m) n += (int64_t)(p->L - m) * p->Jh;
  return n;
}

extern "C" int dino_plan_table(const dino_plan* p, double* out) {
  DINO_REQUIRE(p && out, DINO_EINVAL, "dino_plan_table: NULL argument");
  std::vector<double> buf(p->table_elems);
  DINO_HIP(hipMemcpy(buf.data(), p->p64, sizeof(double) * p->table_elems,
                     hipMemcpyDeviceToHost));
  int64_t src = 0, dst = 0;
  for (int m = 0; m < p->M; ++m) {
    for (int l = m; l < p->L; ++l) {
      std::memcpy(out + dst, buf.data() + src, sizeof(double) * p->Jh);
      dst += p->Jh;
      src += p->Jhp;
    }
  }
  return DINO_OK;
}

extern "C" size_t dino_sht_workspace_bytes(const dino_plan* p, int64_t batch,
                                           int dtype) {
  if (!p || batch < 0) return 0;
  size_t es = dtype == DINO_F64 ? 8 : 4;
  return es * (size_t)batch * p->R * p->J;
}
