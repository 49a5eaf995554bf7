// extern "C" entry points of libkhx.so (declared in include/khx.h).
#include "../../include/khx.h"

#include <string>

#include "khx_internal.h"

namespace khx {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// defined in the other translation units
int kvectors(hipStream_t s, const double* g, int N, const cd* kp, const double* wl, int batch, cd* kx, cd* ky);
int convmat(hipStream_t s, const double* eps, int cplx, int nx, int ny, int P, int Q, int count, cd* C, cd* Fws);
size_t patterned_layer_ws_bytes(int n, int batch);
int patterned_layer(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky, int N,
                    int batch, const double* k0d, cd* S, cd* Wout, cd* Vout, cd* lamout, int* info, void* ws);
int analytic_layer(hipStream_t s, int kind, const cd* kx, const cd* ky, int N, int batch, cd eps,
                   const double* k0d, cd* S, cd* W, cd* V, cd* lam);
size_t star_ws_bytes(int n, int batch);
int star(hipStream_t s, const cd* SA, long long sA, const cd* SB, long long sB, int n, int batch, cd* S,
         int* info, void* ws);
int poynting(hipStream_t s, const cd* S, long long sS, int N, const double* g, const cd* kp, const double* wl,
             const cd* te, const cd* tm, cd epsi, cd epse, int batch, double* R, double* T, double* Rord,
             double* Tord, cd* ws);
int poynting_vec(hipStream_t s, const cd* sR, const cd* sT, int N, const double* g, const cd* kp, const double* wl,
                 cd epsi, cd epse, int batch, double* R, double* T, double* Rord, double* Tord);
int incident(hipStream_t s, int N, const cd* kp, const double* wl, const cd* te, const cd* tm, cd epsi, int batch,
             cd* c);
int zgetrf(hipStream_t s, int n, Mat A, int batch, int* ipiv, int* info);
int zgetrs(hipStream_t s, int n, int m, Mat A, const int* ipiv, Mat B, int batch);
int zdiag_scale(hipStream_t s, Mat A, const cd* dl, const cd* dr, int rows, int cols, int batch, Mat D);
int zidentity_minus(hipStream_t s, const cd* A, int n, int batch, cd* D);
int propagation(hipStream_t s, const cd* kx, const cd* ky, int N, int batch, const double* k0d, cd* X);
int eigenbasis(hipStream_t s, const cd* W, const cd* V, int n, int batch, cd* R);
int mode_propagate(hipStream_t s, const cd* lam, const cd* a, const double* zbar, int n, int nz, cd* M);
int longitudinal(hipStream_t s, const cd* kx, const cd* ky, const cd* F, int N, int nz, cd* uz, cd* w);
int idft_phase(hipStream_t s, const cd* kx, const cd* ky, double k0, const double* x, const double* y, int N,
               int npts, cd* E);

int joint_subspace(hipStream_t s, const cd* in, int Nb, int M, int blocks, int kind, int items, cd* out);
int joint_subspace_vec(hipStream_t s, const cd* in, int Nb, int M, int kind, int items, cd* out);

}  // namespace khx


static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
static inline cd* as_cd(double* p) { return reinterpret_cast<cd*>(p); }
static inline const cd* as_cd(const double* p) { return reinterpret_cast<const cd*>(p); }

#define KHX_ENTRY_BEGIN khx::set_error("");

extern "C" {

const char* khx_last_error(void) { return khx::last_error(); }
int khx_abi_version(void) { return 2; }

int khx_kvectors(const double* g, int N, const double* kp, const double* wl, int batch, double* kx, double* ky,
                 void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0, "khx_kvectors: bad sizes");
  return khx::kvectors(as_stream(stream), g, N, as_cd(kp), wl, batch, as_cd(kx), as_cd(ky));
}

size_t khx_convmat_ws_bytes(int P, int Q, int count) {
  return (size_t)(2 * P - 1) * (2 * Q - 1) * sizeof(cd) * (size_t)(count > 0 ? count : 0) + 256;
}

int khx_convmat(const double* eps, int is_complex, int nx, int ny, int P, int Q, int count, double* C, void* ws,
                void* stream) {
  KHX_ENTRY_BEGIN
  return khx::convmat(as_stream(stream), eps, is_complex, nx, ny, P, Q, count, as_cd(C), static_cast<cd*>(ws));
}

int khx_zgesv(int n, int nrhs, double* A, double* B, int batch, int* info, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && nrhs > 0 && batch >= 0, "khx_zgesv: bad sizes");
  const long long nn = (long long)n * n;
  return khx::zgesv(as_stream(stream), n, nrhs, khx::Mat{as_cd(A), n, nn}, khx::Mat{as_cd(B), nrhs, (long long)n * nrhs},
                    batch, info);
}

int khx_zgemm(int M, int N, int K, const double* alpha, const double* A, int lda, long long strideA,
              const double* B, int ldb, long long strideB, const double* beta, double* C, int ldc,
              long long strideC, int batch, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(M > 0 && N > 0 && K > 0 && batch >= 0, "khx_zgemm: bad sizes");
  KHX_REQUIRE(lda >= K && ldb >= N && ldc >= N, "khx_zgemm: leading dimensions too small");
  const cd al = mkc(alpha[0], alpha[1]), be = mkc(beta[0], beta[1]);
  return khx::zgemm(as_stream(stream), M, N, K, al, khx::Mat{const_cast<cd*>(as_cd(A)), lda, strideA},
                    khx::Mat{const_cast<cd*>(as_cd(B)), ldb, strideB}, be, khx::Mat{as_cd(C), ldc, strideC}, batch);
}

size_t khx_eig_ws_bytes(int n, int batch) { return khx::zeig_workspace_bytes(n, batch) + 256; }

int khx_eig(int n, double* A, int batch, double* lam, double* W, int* info, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_eig: bad sizes");
  const long long nn = (long long)n * n;
  return khx::zeig(as_stream(stream), n, khx::Mat{as_cd(A), n, nn}, as_cd(lam), n, khx::Mat{as_cd(W), n, nn},
                   batch, info, ws);
}

int khx_eig_timed(int n, double* A, int batch, double* lam, double* W, int* info, void* ws,
                  unsigned long long* tphase, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0 && tphase, "khx_eig_timed: bad arguments");
  const long long nn = (long long)n * n;
  return khx::zeig(as_stream(stream), n, khx::Mat{as_cd(A), n, nn}, as_cd(lam), n, khx::Mat{as_cd(W), n, nn},
                   batch, info, ws, tphase);
}

size_t khx_patterned_layer_ws_bytes(int n, int batch) { return khx::patterned_layer_ws_bytes(n, batch); }

int khx_patterned_layer(const double* C, const double* IC, int per_item, const double* kx, const double* ky, int N,
                        int batch, const double* k0d, double* S, double* W, double* V, double* lam, int* info, void* ws,
                        void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0, "khx_patterned_layer: bad sizes");
  return khx::patterned_layer(as_stream(stream), as_cd(C), as_cd(IC), per_item ? (long long)N * N : 0, as_cd(kx),
                              as_cd(ky), N, batch, k0d, as_cd(S), as_cd(W), as_cd(V), as_cd(lam), info, ws);
}

int khx_uniform_layer(const double* kx, const double* ky, int N, int batch, const double* eps_host,
                      const double* k0d, double* S, double* W, double* V, double* lam, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && eps_host && k0d, "khx_uniform_layer: bad arguments");
  return khx::analytic_layer(as_stream(stream), 0, as_cd(kx), as_cd(ky), N, batch,
                             mkc(eps_host[0], eps_host[1]), k0d, as_cd(S), as_cd(W), as_cd(V), as_cd(lam));
}

int khx_halfspace(int side, const double* kx, const double* ky, int N, int batch, const double* eps_host,
                  double* S, double* W, double* V, double* lam, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(side == 1 || side == 2, "khx_halfspace: side must be 1 (reflection) or 2 (transmission)");
  KHX_REQUIRE(N > 0 && batch >= 0 && eps_host, "khx_halfspace: bad arguments");
  return khx::analytic_layer(as_stream(stream), side, as_cd(kx), as_cd(ky), N, batch,
                             mkc(eps_host[0], eps_host[1]), nullptr, as_cd(S), as_cd(W), as_cd(V), as_cd(lam));
}

size_t khx_redheffer_ws_bytes(int n, int batch) { return khx::star_ws_bytes(n, batch); }

int khx_redheffer(const double* SA, int a_bcast, const double* SB, int b_bcast, int n, int batch, double* S,
                  int* info, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_redheffer: bad sizes");
  KHX_REQUIRE(S != SA && S != SB, "khx_redheffer: S must not alias an operand");
  const long long s4 = 4ll * n * n;
  return khx::star(as_stream(stream), as_cd(SA), a_bcast ? 0 : s4, as_cd(SB), b_bcast ? 0 : s4, n, batch, as_cd(S),
                   info, ws);
}

size_t khx_poynting_ws_bytes(int N, int batch) { return (size_t)3 * 2 * N * sizeof(cd) * batch + 256; }

int khx_poynting(const double* Stot, int N, const double* g, const double* kp, const double* wl, const double* te,
                 const double* tm, const double* epsi_host, const double* epse_host, int batch, double* R, double* T,
                 double* Rord, double* Tord, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && epsi_host && epse_host && ws, "khx_poynting: bad arguments");
  const long long s4 = 4ll * (2 * N) * (2 * N);
  return khx::poynting(as_stream(stream), as_cd(Stot), s4, N, g, as_cd(kp), wl, as_cd(te), as_cd(tm),
                       mkc(epsi_host[0], epsi_host[1]), mkc(epse_host[0], epse_host[1]), batch, R, T, Rord,
                       Tord, static_cast<cd*>(ws));
}

int khx_incident(int N, const double* kp, const double* wl, const double* te, const double* tm,
                 const double* epsi_host, int batch, double* c, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && epsi_host, "khx_incident: bad arguments");
  return khx::incident(as_stream(stream), N, as_cd(kp), wl, as_cd(te), as_cd(tm), mkc(epsi_host[0], epsi_host[1]),
                       batch, as_cd(c));
}

int khx_poynting_vec(const double* sR, const double* sT, int N, const double* g, const double* kp, const double* wl,
                     const double* epsi_host, const double* epse_host, int batch, double* R, double* T, double* Rord,
                     double* Tord, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && epsi_host && epse_host, "khx_poynting_vec: bad arguments");
  return khx::poynting_vec(as_stream(stream), as_cd(sR), as_cd(sT), N, g, as_cd(kp), wl,
                           mkc(epsi_host[0], epsi_host[1]), mkc(epse_host[0], epse_host[1]), batch, R, T, Rord, Tord);
}

int khx_lu_factor(int n, double* A, int batch, int* ipiv, int* info, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_lu_factor: bad sizes");
  return khx::zgetrf(as_stream(stream), n, khx::Mat{as_cd(A), n, (long long)n * n}, batch, ipiv, info);
}

int khx_lu_solve(int n, int nrhs, const double* LU, const int* ipiv, double* B, int batch, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && nrhs > 0 && batch >= 0, "khx_lu_solve: bad sizes");
  return khx::zgetrs(as_stream(stream), n, nrhs, khx::Mat{const_cast<cd*>(as_cd(LU)), n, (long long)n * n}, ipiv,
                     khx::Mat{as_cd(B), nrhs, (long long)n * nrhs}, batch);
}

int khx_diag_scale(const double* A, int lda, long long strideA, const double* dl, const double* dr, int rows,
                   int cols, int batch, double* out, int ldo, long long strideO, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(rows > 0 && cols > 0 && batch >= 0 && lda >= cols && ldo >= cols, "khx_diag_scale: bad sizes");
  return khx::zdiag_scale(as_stream(stream), khx::Mat{const_cast<cd*>(as_cd(A)), lda, strideA}, as_cd(dl),
                          as_cd(dr), rows, cols, batch, khx::Mat{as_cd(out), ldo, strideO});
}

int khx_propagation(const double* kx, const double* ky, int N, int batch, const double* k0d, double* X,
                    void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0, "khx_propagation: bad sizes");
  return khx::propagation(as_stream(stream), as_cd(kx), as_cd(ky), N, batch, k0d, as_cd(X));
}

int khx_eigenbasis(const double* W, const double* V, int n, int batch, double* R, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_eigenbasis: bad sizes");
  return khx::eigenbasis(as_stream(stream), as_cd(W), as_cd(V), n, batch, as_cd(R));
}

int khx_mode_propagate(const double* lam, const double* a, const double* zbar, int n, int nz, double* M,
                       void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && nz >= 0, "khx_mode_propagate: bad sizes");
  return khx::mode_propagate(as_stream(stream), as_cd(lam), as_cd(a), zbar, n, nz, as_cd(M));
}

int khx_longitudinal(const double* kx, const double* ky, const double* F, int N, int nz, double* uz, double* w,
                     void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && nz >= 0, "khx_longitudinal: bad sizes");
  return khx::longitudinal(as_stream(stream), as_cd(kx), as_cd(ky), as_cd(F), N, nz, as_cd(uz), as_cd(w));
}

int khx_idft_phase(const double* kx, const double* ky, double k0, const double* x, const double* y, int N, int npts,
                   double* E, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && npts >= 0, "khx_idft_phase: bad sizes");
  return khx::idft_phase(as_stream(stream), as_cd(kx), as_cd(ky), k0, x, y, N, npts, as_cd(E));
}

int khx_identity_minus(const double* A, int n, int batch, double* out, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_identity_minus: bad sizes");
  return khx::zidentity_minus(as_stream(stream), as_cd(A), n, batch, as_cd(out));
}

int khx_joint_subspace(const double* in, int Nb, int blocks, int kind, int items, double* out, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(Nb > 0 && (blocks == 1 || blocks == 4) && items >= 0, "khx_joint_subspace: bad arguments");
  return khx::joint_subspace(as_stream(stream), as_cd(in), Nb, Nb, blocks, kind, items, as_cd(out));
}

int khx_joint_subspace_diag(const double* in, int Nb, int kind, int items, double* out, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(Nb > 0 && items >= 0, "khx_joint_subspace_diag: bad arguments");
  return khx::joint_subspace_vec(as_stream(stream), as_cd(in), Nb, Nb, kind, items, as_cd(out));
}

}  // extern "C"
