// Semi-implicit operator: creation, argument checks and C entry points.
//
// The reference inverts, per total wavenumber l, the (2K+1)^2 block matrix
//   [[I, eta lam_l G, eta R lam_l T_ref], [eta H, I, 0], [eta dsigma^T, 0, 1]]
// on the host (primitive_equations.py:544-598, 1273-1330).  Eliminating the
// temperature and log-pressure rows leaves, for the divergence,
//   (I - eta^2 lam_l A) delta'' = delta - eta lam_l (G T + R T_ref lnps),
//   A = G H + R T_ref dsigma^T   (independent of l),
// which the caller factorises once in FP64 as A = V diag(d) V^{-1}.  The
// kernels (implicit_mfma.hip) apply the resulting l-independent matrices to
// 16-column batches on the FP64 matrix cores.
#include <cstring>
#include <vector>

#include "dino_common.h"

namespace dino {

template <typename T>
int implicit_mfma_dispatch(bool solve, bool adjoint, bool acc, const dino_implicit* f,
                           const void* div, const void* temp, const void* lnps, void* d_o,
                           void* t_o, void* l_o, int64_t members, hipStream_t s);
int implicit_mfma_tables(int K, const double* G, const double* H, const double* tref,
                         const double* dsigma, const double* V, const double* Vinv,
                         const double* d, double r_gas, std::vector<double>* out);

namespace {

constexpr int kTables = 15;

int run(bool solve, bool adjoint, bool acc, const dino_implicit* f, const void* div,
        const void* temp, const void* lnps, void* d_o, void* t_o, void* l_o, int64_t members,
        int dtype, void* stream) {
  if (dtype == DINO_F64)
    return implicit_mfma_dispatch<double>(solve, adjoint, acc, f, div, temp, lnps, d_o, t_o,
                                          l_o, members, as_stream(stream));
  return implicit_mfma_dispatch<float>(solve, adjoint, acc, f, div, temp, lnps, d_o, t_o, l_o,
                                       members, as_stream(stream));
}

void free_implicit(dino_implicit* f) {
  if (!f) return;
  void* bufs[] = {f->G, f->H, f->V, f->Vinv, f->tref, f->dsigma, f->d, f->lam};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (double* b : f->mtab)
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
  std::vector<double> tabs[kTables];
  implicit_mfma_tables(K, G, H, tref, dsigma, V, Vinv, d, r_gas, tabs);
  for (int i = 0; i < kTables; ++i) {
    hipError_t e = hipMalloc(&f->mtab[i], sizeof(double) * tabs[i].size());
    if (e == hipSuccess)
      e = hipMemcpy(f->mtab[i], tabs[i].data(), sizeof(double) * tabs[i].size(),
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      free_implicit(f);
      return hip_fail(e, "implicit fragment tables");
    }
  }
  *out = f;
  return DINO_OK;
}

extern "C" int dino_implicit_destroy(dino_implicit* f) {
  free_implicit(f);
  return DINO_OK;
}

static int checked(const char* name, bool solve, bool adjoint, const dino_implicit* f,
                   const void* div, const void* temp, const void* lnps, void* div_out,
                   void* temp_out, void* lnps_out, int64_t members, int dtype, void* stream,
                   bool acc = false) {
  DINO_REQUIRE(f && div && temp && lnps && div_out && temp_out && lnps_out, DINO_EINVAL,
               "%s: NULL argument", name);
  DINO_REQUIRE(!solve || (f->V && f->Vinv && f->d), DINO_EINVAL,
               "%s: factor has no vertical modes", name);
  DINO_REQUIRE(members >= 0, DINO_EINVAL, "%s: negative member count", name);
  DINO_REQUIRE(dtype == DINO_F32 || dtype == DINO_F64, DINO_EINVAL, "%s: invalid dtype %d",
               name, dtype);
  return run(solve, adjoint, acc, f, div, temp, lnps, div_out, temp_out, lnps_out, members,
             dtype, stream);
}

extern "C" int dino_implicit_solve(const dino_implicit* f, const void* div, const void* temp,
                                   const void* lnps, void* div_out, void* temp_out,
                                   void* lnps_out, int64_t members, int dtype, void* stream) {
  return checked("dino_implicit_solve", true, false, f, div, temp, lnps, div_out, temp_out,
                 lnps_out, members, dtype, stream);
}

extern "C" int dino_implicit_terms(const dino_implicit* f, const void* div, const void* temp,
                                   const void* lnps, void* div_out, void* temp_out,
                                   void* lnps_out, int64_t members, int dtype, void* stream) {
  return checked("dino_implicit_terms", false, false, f, div, temp, lnps, div_out, temp_out,
                 lnps_out, members, dtype, stream);
}

extern "C" int dino_implicit_solve_adjoint(const dino_implicit* f, const void* div_cot,
                                           const void* temp_cot, const void* lnps_cot,
                                           void* div_out, void* temp_out, void* lnps_out,
                                           int64_t members, int dtype, void* stream) {
  return checked("dino_implicit_solve_adjoint", true, true, f, div_cot, temp_cot, lnps_cot,
                 div_out, temp_out, lnps_out, members, dtype, stream);
}

extern "C" int dino_implicit_terms_adjoint(const dino_implicit* f, const void* div_cot,
                                           const void* temp_cot, const void* lnps_cot,
                                           void* div_out, void* temp_out, void* lnps_out,
                                           int64_t members, int dtype, void* stream) {
  return checked("dino_implicit_terms_adjoint", false, true, f, div_cot, temp_cot, lnps_cot,
                 div_out, temp_out, lnps_out, members, dtype, stream);
}

extern "C" int dino_implicit_terms_adjoint_accumulate(
    const dino_implicit* f, const void* div_cot, const void* temp_cot, const void* lnps_cot,
    void* div_out, void* temp_out, void* lnps_out, int64_t members, int dtype, void* stream) {
  return checked("dino_implicit_terms_adjoint_accumulate", false, true, f, div_cot, temp_cot,
                 lnps_cot, div_out, temp_out, lnps_out, members, dtype, stream, true);
}
