// Semi-implicit solve in FP64 through vertical modes, and implicit terms.
//
// The reference inverts, per total wavenumber l, the (2K+1)^2 block matrix
//   [[I, eta lam_l G, eta R lam_l T_ref], [eta H, I, 0], [eta dsigma^T, 0, 1]]
// on the host (primitive_equations.py:544-598, 1273-1330).  Eliminating the
// temperature and log-pressure rows leaves, for the divergence,
//   (I - eta^2 lam_l A) delta'' = delta - eta lam_l (G T + R T_ref lnps),
//   A = G H + R T_ref dsigma^T   (independent of l),
// and with A = V diag(d) V^{-1} (factorised once in FP64 by the caller):
//   delta'' = V diag(1 / (1 - eta^2 lam_l d)) V^{-1} rhs,
//   T''     = T - eta H delta'',   lnps'' = lnps - eta dsigma . delta''.
// Every matrix is l-independent, so one thread per (member, m-row, l) column
// reads its 2K+1 values with coalesced loads (adjacent threads = adjacent l)
// and takes the matrices through uniform (scalar-cache) loads.
#include <cstring>

#include "dino_common.h"

namespace dino {
namespace {

constexpr int kImpThreads = 128;

// Per-thread vectors live in registers while they are the operand of a
// matvec (inner loop unrolled over KMAX) and in LDS (column tid) between
// matvecs, so only the inner dimension is unrolled.
template <typename T, int KMAX>
__global__ __launch_bounds__(kImpThreads) void implicit_solve_kernel(
    const T* __restrict__ div, const T* __restrict__ temp, const T* __restrict__ lnps,
    T* __restrict__ div_out, T* __restrict__ temp_out, T* __restrict__ lnps_out,
    const double* __restrict__ G, const double* __restrict__ H,
    const double* __restrict__ V, const double* __restrict__ Vinv,
    const double* __restrict__ tref, const double* __restrict__ dsig,
    const double* __restrict__ d, const double* __restrict__ lam, int K, int L,
    int64_t cols, int64_t members, double r_gas, double eta) {
  __shared__ double buf[KMAX][kImpThreads];
  const int tid = threadIdx.x;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + tid;
  if (gid >= cols * members) return;  // no block-wide barriers below
  int64_t b = gid / cols, c = gid % cols;
  int l = (int)(c % L);
  const double lam_l = lam[l];
  const int64_t base = b * (int64_t)K * cols + c;
  double x[KMAX];
  const double lp = (double)lnps[b * cols + c];
#pragma unroll
  for (int h = 0; h < KMAX; ++h) x[h] = h < K ? (double)temp[base + h * cols] : 0.0;
  // rhs = delta - eta lam (G T + R T_ref lnps)
  for (int g = 0; g < K; ++g) {
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < KMAX; ++h)
      if (h < K) acc += G[g * K + h] * x[h];
    acc += r_gas * tref[g] * lp;
    buf[g][tid] = (double)div[base + g * cols] - eta * lam_l * acc;
  }
#pragma unroll
  for (int h = 0; h < KMAX; ++h) x[h] = h < K ? buf[h][tid] : 0.0;
  // w = diag(1 / (1 - eta^2 lam d)) Vinv rhs
  for (int i = 0; i < K; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int g = 0; g < KMAX; ++g)
      if (g < K) acc += Vinv[i * K + g] * x[g];
    buf[i][tid] = acc / (1.0 - eta * eta * lam_l * d[i]);
  }
#pragma unroll
  for (int h = 0; h < KMAX; ++h) x[h] = h < K ? buf[h][tid] : 0.0;
  // delta'' = V w
  for (int g = 0; g < K; ++g) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < KMAX; ++i)
      if (i < K) acc += V[g * K + i] * x[i];
    buf[g][tid] = acc;
    div_out[base + g * cols] = (T)acc;
  }
#pragma unroll
  for (int h = 0; h < KMAX; ++h) x[h] = h < K ? buf[h][tid] : 0.0;
  // T'' = T - eta H delta'' ; lnps'' = lnps - eta dsigma . delta''
  for (int g = 0; g < K; ++g) {
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < KMAX; ++h)
      if (h < K) acc += H[g * K + h] * x[h];
    temp_out[base + g * cols] = (T)((double)temp[base + g * cols] - eta * acc);
  }
  double sacc = 0.0;
#pragma unroll
  for (int h = 0; h < KMAX; ++h)
    if (h < K) sacc += dsig[h] * x[h];
  lnps_out[b * cols + c] = (T)(lp - eta * sacc);
}

template <typename T, int KMAX>
__global__ __launch_bounds__(kImpThreads) void implicit_terms_kernel(
    const T* __restrict__ div, const T* __restrict__ temp, const T* __restrict__ lnps,
    T* __restrict__ div_out, T* __restrict__ temp_out, T* __restrict__ lnps_out,
    const double* __restrict__ G, const double* __restrict__ H,
    const double* __restrict__ tref, const double* __restrict__ dsig,
    const double* __restrict__ lam, int K, int L, int64_t cols, int64_t members,
    double r_gas) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= cols * members) return;
  int64_t b = gid / cols, c = gid % cols;
  int l = (int)(c % L);
  const double lam_l = lam[l];
  const int64_t base = b * (int64_t)K * cols + c;
  double x[KMAX];
  const double lp = (double)lnps[b * cols + c];
#pragma unroll
  for (int h = 0; h < KMAX; ++h) x[h] = h < K ? (double)temp[base + h * cols] : 0.0;
  for (int g = 0; g < K; ++g) {
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < KMAX; ++h)
      if (h < K) acc += G[g * K + h] * x[h];
    div_out[base + g * cols] = (T)(-lam_l * (acc + r_gas * tref[g] * lp));
  }
#pragma unroll
  for (int h = 0; h < KMAX; ++h) x[h] = h < K ? (double)div[base + h * cols] : 0.0;
  double sacc = 0.0;
  for (int g = 0; g < K; ++g) {
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < KMAX; ++h)
      if (h < K) acc += H[g * K + h] * x[h];
    temp_out[base + g * cols] = (T)(-acc);
  }
#pragma unroll
  for (int h = 0; h < KMAX; ++h)
    if (h < K) sacc += dsig[h] * x[h];
  lnps_out[b * cols + c] = (T)(-sacc);
}

template <typename T, int KMAX>
int launch_k(bool solve, const dino_implicit* f, const void* div, const void* temp,
             const void* lnps, void* d_o, void* t_o, void* l_o, int64_t members,
             hipStream_t s) {
  int64_t cols = (int64_t)f->R * f->L;
  int64_t n = cols * members;
  if (n == 0) return DINO_OK;
  unsigned blocks = (unsigned)((n + kImpThreads - 1) / kImpThreads);
  if (solve) {
    implicit_solve_kernel<T, KMAX><<<blocks, kImpThreads, 0, s>>>(
        (const T*)div, (const T*)temp, (const T*)lnps, (T*)d_o, (T*)t_o, (T*)l_o, f->G, f->H,
        f->V, f->Vinv, f->tref, f->dsigma, f->d, f->lam, f->K, f->L, cols, members, f->r_gas,
        f->eta);
  } else {
    implicit_terms_kernel<T, KMAX><<<blocks, kImpThreads, 0, s>>>(
        (const T*)div, (const T*)temp, (const T*)lnps, (T*)d_o, (T*)t_o, (T*)l_o, f->G, f->H,
        f->tref, f->dsigma, f->lam, f->K, f->L, cols, members, f->r_gas);
  }
  DINO_LAUNCH_CHECK("implicit kernel");
  return DINO_OK;
}

template <typename T>
int dispatch(bool solve, const dino_implicit* f, const void* div, const void* temp,
             const void* lnps, void* d_o, void* t_o, void* l_o, int64_t members, hipStream_t s) {
  int K = f->K;
  if (K <= 16) return launch_k<T, 16>(solve, f, div, temp, lnps, d_o, t_o, l_o, members, s);
  if (K <= 32) return launch_k<T, 32>(solve, f, div, temp, lnps, d_o, t_o, l_o, members, s);
  if (K <= 40) return launch_k<T, 40>(solve, f, div, temp, lnps, d_o, t_o, l_o, members, s);
  if (K <= 48) return launch_k<T, 48>(solve, f, div, temp, lnps, d_o, t_o, l_o, members, s);
  return launch_k<T, 64>(solve, f, div, temp, lnps, d_o, t_o, l_o, members, s);
}

void free_implicit(dino_implicit* f) {
  if (!f) return;
  void* bufs[] = {f->G, f->H, f->V, f->Vinv, f->tref, f->dsigma, f->d, f->lam};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete f;
}

}  // namespace
}  // namespace dino

using namespace dino;

extern "C" int dino_implicit_create(int K, int L, int R, const double* G, const double* H,
                                    const double* tref, const double* dsigma,
                                    const double* lam, const double* V, const double* Vinv,
                                    const double* d, double r_gas, double eta,
                                    dino_implicit** out) {
  DINO_REQUIRE(out, DINO_EINVAL, "dino_implicit_create: out is NULL");
  *out = nullptr;
  DINO_REQUIRE(K >= 1 && K <= 64, DINO_EINVAL,
               "implicit solve supports 1..64 levels, got %d", K);
  DINO_REQUIRE(L >= 1 && R >= 1, DINO_EINVAL, "invalid modal shape (%d, %d)", R, L);
  DINO_REQUIRE(G && H && tref && dsigma && lam, DINO_EINVAL, "NULL matrix argument");
  dino_implicit* f = new dino_implicit();
  std::memset(f, 0, sizeof(*f));
  f->K = K;
  f->L = L;
  f->R = R;
  f->r_gas = r_gas;
  f->eta = eta;
  struct Item {
    double** dst;
    const double* src;
    size_t n;
  } items[] = {
      {&f->G, G, (size_t)K * K},       {&f->H, H, (size_t)K * K},
      {&f->V, V, (size_t)K * K},       {&f->Vinv, Vinv, (size_t)K * K},
      {&f->tref, tref, (size_t)K},     {&f->dsigma, dsigma, (size_t)K},
      {&f->d, d, (size_t)K},           {&f->lam, lam, (size_t)L},
  };
  for (auto& it : items) {
    hipError_t e = hipMalloc(it.dst, sizeof(double) * it.n);
    if (e != hipSuccess) {
      free_implicit(f);
      return hip_fail(e, "hipMalloc(implicit)");
    }
    if (it.src) {
      e = hipMemcpy(*it.dst, it.src, sizeof(double) * it.n, hipMemcpyHostToDevice);
    } else {
      e = hipMemset(*it.dst, 0, sizeof(double) * it.n);
    }
    if (e != hipSuccess) {
      free_implicit(f);
      return hip_fail(e, "hipMemcpy(implicit)");
    }
  }
  *out = f;
  return DINO_OK;
}

extern "C" int dino_implicit_destroy(dino_implicit* f) {
  free_implicit(f);
  return DINO_OK;
}

extern "C" int dino_implicit_solve(const dino_implicit* f, const void* div, const void* temp,
                                   const void* lnps, void* div_out, void* temp_out,
                                   void* lnps_out, int64_t members, int dtype, void* stream) {
  DINO_REQUIRE(f && div && temp && lnps && div_out && temp_out && lnps_out, DINO_EINVAL,
               "dino_implicit_solve: NULL argument");
  DINO_REQUIRE(f->V && f->Vinv && f->d, DINO_EINVAL,
               "dino_implicit_solve: factor has no vertical modes");
  DINO_REQUIRE(members >= 0, DINO_EINVAL, "negative member count");
  if (dtype == DINO_F64)
    return dispatch<double>(true, f, div, temp, lnps, div_out, temp_out, lnps_out, members,
                            as_stream(stream));
  return dispatch<float>(true, f, div, temp, lnps, div_out, temp_out, lnps_out, members,
                         as_stream(stream));
}

extern "C" int dino_implicit_terms(const dino_implicit* f, const void* div, const void* temp,
                                   const void* lnps, void* div_out, void* temp_out,
                                   void* lnps_out, int64_t members, int dtype, void* stream) {
  DINO_REQUIRE(f && div && temp && lnps && div_out && temp_out && lnps_out, DINO_EINVAL,
               "dino_implicit_terms: NULL argument");
  DINO_REQUIRE(members >= 0, DINO_EINVAL, "negative member count");
  if (dtype == DINO_F64)
    return dispatch<double>(false, f, div, temp, lnps, div_out, temp_out, lnps_out, members,
                            as_stream(stream));
  return dispatch<float>(false, f, div, temp, lnps, div_out, temp_out, lnps_out, members,
                         as_stream(stream));
}
