// extern "C" entry points of libkhx.so (declared in include/khx.h).
#include "../../include/khx.h"

#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "khx_internal.h"

namespace khx {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

int ensure_dynamic_lds(const void* fn, size_t bytes) {
  if (bytes <= 64 * 1024) return 0;
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> set_bytes;  // (device, kernel) -> attribute value
  int dev = 0;
  KHX_CHECK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  size_t& cur = set_bytes[std::make_pair(dev, fn)];
  if (bytes <= cur) return 0;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipFuncSetAttribute(MaxDynamicSharedMemorySize): ") + hipGetErrorString(e));
    return -3;
  }
  cur = bytes;
  return 0;
}

// defined in the other translation units
size_t patterned_layer_bwd_ws_bytes(int n, int batch);
int patterned_layer_bwd(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky,
                        int N, int batch, const double* k0d, const cd* W, const cd* V, const cd* lam, const cd* Sbar,
                        double eta_rel, cd* Cbar, int* info, void* ws, const cd* cache);
size_t star_bwd_ws_bytes(int n, int batch);
int star_bwd(hipStream_t s, const cd* SA, long long sA, const cd* SB, long long sB, int n, int batch, const cd* Sbar,
             cd* SAbar, cd* SBbar, int* info, void* ws);
size_t convmat_adjoint_ws_bytes(int P, int Q, int batch);
int convmat_adjoint(hipStream_t s, const cd* Cbar, int P, int Q, int nx, int ny, int batch, int cplx,
                    double* epsbar, cd* Fws);
int poynting_vec_adjoint(hipStream_t s, const cd* sR, const cd* sT, int N, const double* g, const cd* kp,
                         const double* wl, cd epsi, cd epse, const double* wR, const double* wT, int batch,
                         cd* sRbar, cd* sTbar);
int kvectors(hipStream_t s, const double* g, int N, const cd* kp, const double* wl, int batch, cd* kx, cd* ky);
int convmat(hipStream_t s, const double* eps, int cplx, int nx, int ny, int P, int Q, int count, cd* C, cd* Fws);
size_t patterned_layer_ws_bytes(int n, int batch);
int patterned_layer(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky, int N,
                    int batch, const double* k0d, cd* S, cd* Wout, cd* Vout, cd* lamout, int* info, void* ws,
                    cd* cache);
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

int khx_beam_fields(int kind, const double* x, const double* y, const double* z, long long npts, double wl,
                    double w0, double er, int p, int l, const double* geom, double* E, double* H, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(x && y && z && geom && E && H, "khx_beam_fields: null argument");
  return khx::beam_fields(as_stream(stream), kind, x, y, z, npts, wl, w0, er, p, l, geom, as_cd(E), as_cd(H));
}

int khx_bloch_amplitudes(const double* fields, const double* x, const double* y, int ns, int bz0, int bz1,
                         double kpx, double kpy, const double* gx, const double* gy, int nk, double* out,
                         void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(fields && x && y && gx && gy && out, "khx_bloch_amplitudes: null argument");
  return khx::bloch_amplitudes(as_stream(stream), reinterpret_cast<const cd*>(fields), x, y, ns, bz0, bz1, kpx, kpy,
                               gx, gy, nk, as_cd(out));
}

int khx_points_dft(const double* f, const double* x, const double* y, long long npts, const double* kx,
                   const double* ky, int nk, int batch, double scale, double* out, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(f && x && y && kx && ky && out, "khx_points_dft: null argument");
  return khx::points_dft(as_stream(stream), reinterpret_cast<const cd*>(f), x, y, npts, kx, ky, nk, batch, scale,
                         as_cd(out));
}

int khx_scattering_splitlr(int ng, const double* S, int batch, double* Sl, double* Sr, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(S && Sl && Sr, "khx_scattering_splitlr: null argument");
  return khx::scattering_splitlr(as_stream(stream), ng, reinterpret_cast<const cd*>(S), batch, as_cd(Sl), as_cd(Sr));
}

size_t khx_scattering_eig_ws_bytes(int ng, int batch) {
  return ng < 1 || batch < 0 ? 0 : khx::scattering_eig_ws_bytes(ng, batch);
}

int khx_scattering_eig(int ng, const double* S, int batch, double sigma_re, double sigma_im, double* w, double* v,
                       int* info, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(S && w && v && info && ws, "khx_scattering_eig: null argument");
  return khx::scattering_eig(as_stream(stream), ng, reinterpret_cast<const cd*>(S), batch, mkc(sigma_re, sigma_im),
                             as_cd(w), as_cd(v), info, ws);
}

size_t khx_scattering_det_ws_bytes(int ng, int batch) {
  return ng < 1 || batch < 0 ? 0 : khx::scattering_det_ws_bytes(ng, batch);
}

int khx_scattering_det(int ng, const double* S, int batch, double* det, int* info, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(S && det && info && ws, "khx_scattering_det: null argument");
  return khx::scattering_det(as_stream(stream), ng, reinterpret_cast<const cd*>(S), batch, as_cd(det), info, ws);
}

int khx_convmat_islands(const int* kinds, const double* params, const double* eps, int nisl,
                        const double* eps_host, const double* recip, double sigma, int P, int Q, double* C, void* ws,
                        void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(eps_host && recip && C && ws && (nisl == 0 || (kinds && params && eps)),
              "khx_convmat_islands: null argument");
  return khx::convmat_islands(as_stream(stream), kinds, params, reinterpret_cast<const cd*>(eps), nisl,
                              mkc(eps_host[0], eps_host[1]), recip, sigma, P, Q, as_cd(C), static_cast<cd*>(ws));
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

int khx_eig_noalloc(int n, double* A, int batch, double* lam, double* W, int* info, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_eig_noalloc: bad sizes");
  const long long nn = (long long)n * n;
  return khx::zeig_noalloc(as_stream(stream), n, khx::Mat{as_cd(A), n, nn}, as_cd(lam), n,
                           khx::Mat{as_cd(W), n, nn}, batch, info);
}

size_t khx_eig_noalloc_bytes(void) { return khx::zeig_noalloc_limit(); }

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
                              as_cd(ky), N, batch, k0d, as_cd(S), as_cd(W), as_cd(V), as_cd(lam), info, ws, nullptr);
}

int khx_patterned_layer_cached(const double* C, const double* IC, int per_item, const double* kx, const double* ky,
                               int N, int batch, const double* k0d, double* S, double* W, double* V, double* lam,
                               int* info, double* cache, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && cache, "khx_patterned_layer_cached: bad arguments");
  return khx::patterned_layer(as_stream(stream), as_cd(C), as_cd(IC), per_item ? (long long)N * N : 0, as_cd(kx),
                              as_cd(ky), N, batch, k0d, as_cd(S), as_cd(W), as_cd(V), as_cd(lam), info, ws,
                              as_cd(cache));
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

// ------------------------------------------------------------------ gradient
int khx_zgemm_op(int opa, int opb, int M, int N, int K, const double* alpha, const double* A, int lda,
                 long long strideA, const double* B, int ldb, long long strideB, const double* beta, double* C,
                 int ldc, long long strideC, int batch, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(M > 0 && N > 0 && K > 0 && batch >= 0, "khx_zgemm_op: bad sizes");
  KHX_REQUIRE(lda >= (opa ? M : K) && ldb >= (opb ? K : N) && ldc >= N, "khx_zgemm_op: leading dimensions too small");
  const cd al = mkc(alpha[0], alpha[1]), be = mkc(beta[0], beta[1]);
  return khx::zgemm_op(as_stream(stream), opa, opb, M, N, K, al, khx::Mat{const_cast<cd*>(as_cd(A)), lda, strideA},
                       khx::Mat{const_cast<cd*>(as_cd(B)), ldb, strideB}, be, khx::Mat{as_cd(C), ldc, strideC},
                       batch);
}

int khx_conj_transpose(const double* A, int lda, long long strideA, int rows, int cols, double* B, int ldb,
                       long long strideB, int batch, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(rows >= 0 && cols >= 0 && batch >= 0 && lda >= cols && ldb >= rows, "khx_conj_transpose: bad sizes");
  return khx::conj_transpose(as_stream(stream), khx::Mat{const_cast<cd*>(as_cd(A)), lda, strideA}, rows, cols,
                             khx::Mat{as_cd(B), ldb, strideB}, batch);
}

int khx_reduce_dot(const double* A, int lda, long long strideA, const double* B, int ldb, long long strideB,
                   int rows, int cols, int axis, int accumulate, double* out, long long ostride, int batch,
                   void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(rows >= 0 && cols >= 0 && batch >= 0 && lda >= cols && ldb >= cols && (axis == 0 || axis == 1),
              "khx_reduce_dot: bad arguments");
  return khx::reduce_dot(as_stream(stream), khx::Mat{const_cast<cd*>(as_cd(A)), lda, strideA},
                         khx::Mat{const_cast<cd*>(as_cd(B)), ldb, strideB}, rows, cols, axis, accumulate,
                         as_cd(out), ostride, batch);
}

int khx_eig_adjoint_mix(double* G, const double* mu, const double* mubar, int n, int batch, double eta_rel,
                        void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0 && eta_rel >= 0.0, "khx_eig_adjoint_mix: bad arguments");
  return khx::eig_adjoint_mix(as_stream(stream), as_cd(G), as_cd(mu), as_cd(mubar), n, batch, eta_rel);
}

int khx_lambda_adjoint(const double* lam, const double* k0d, const double* xbar, const double* ilbar, int n,
                       int batch, double* mubar, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0 && (xbar == nullptr || k0d != nullptr), "khx_lambda_adjoint: bad arguments");
  return khx::lambda_adjoint(as_stream(stream), as_cd(lam), k0d, as_cd(xbar), as_cd(ilbar), n, batch, as_cd(mubar));
}

int khx_build_pq(const double* C, const double* IC, int per_item, const double* kx, const double* ky, int N,
                 int batch, double* P, double* Q, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0, "khx_build_pq: bad sizes");
  return khx::build_pq(as_stream(stream), as_cd(C), as_cd(IC), per_item ? (long long)N * N : 0, as_cd(kx), as_cd(ky),
                       N, batch, as_cd(P), as_cd(Q));
}

int khx_pq_adjoint(const double* Pbar, const double* Qbar, const double* kx, const double* ky, int N, int batch,
                   double* ICbar, double* Cbar, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0, "khx_pq_adjoint: bad sizes");
  return khx::pq_adjoint(as_stream(stream), as_cd(Pbar), as_cd(Qbar), as_cd(kx), as_cd(ky), N, batch, as_cd(ICbar),
                         as_cd(Cbar));
}

size_t khx_convmat_adjoint_ws_bytes(int P, int Q, int batch) { return khx::convmat_adjoint_ws_bytes(P, Q, batch); }

int khx_convmat_adjoint(const double* Cbar, int P, int Q, int nx, int ny, int batch, int is_complex,
                        double* epsbar, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(batch >= 0 && ws != nullptr, "khx_convmat_adjoint: bad arguments");
  return khx::convmat_adjoint(as_stream(stream), as_cd(Cbar), P, Q, nx, ny, batch, is_complex, epsbar,
                              static_cast<cd*>(ws));
}

size_t khx_patterned_layer_bwd_ws_bytes(int n, int batch) { return khx::patterned_layer_bwd_ws_bytes(n, batch); }

int khx_patterned_layer_bwd(const double* C, const double* IC, int per_item, const double* kx, const double* ky,
                            int N, int batch, const double* k0d, const double* W, const double* V,
                            const double* lam, const double* Sbar, double eta_rel, double* Cbar, int* info, void* ws,
                            void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && ws && W && V && lam && Sbar && Cbar && eta_rel >= 0.0,
              "khx_patterned_layer_bwd: bad arguments");
  KHX_REQUIRE(batch <= 65535, "khx_patterned_layer_bwd: at most 65535 items per call");
  return khx::patterned_layer_bwd(as_stream(stream), as_cd(C), as_cd(IC), per_item ? (long long)N * N : 0, as_cd(kx),
                                  as_cd(ky), N, batch, k0d, as_cd(W), as_cd(V), as_cd(lam), as_cd(Sbar), eta_rel,
                                  as_cd(Cbar), info, ws, nullptr);
}

int khx_patterned_layer_bwd_cached(const double* C, const double* IC, int per_item, const double* kx,
                                   const double* ky, int N, int batch, const double* k0d, const double* W,
                                   const double* V, const double* lam, const double* cache, const double* Sbar,
                                   double eta_rel, double* Cbar, int* info, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && ws && W && V && lam && cache && Sbar && Cbar && eta_rel >= 0.0,
              "khx_patterned_layer_bwd_cached: bad arguments");
  KHX_REQUIRE(batch <= 65535, "khx_patterned_layer_bwd_cached: at most 65535 items per call");
  return khx::patterned_layer_bwd(as_stream(stream), as_cd(C), as_cd(IC), per_item ? (long long)N * N : 0, as_cd(kx),
                                  as_cd(ky), N, batch, k0d, as_cd(W), as_cd(V), as_cd(lam), as_cd(Sbar), eta_rel,
                                  as_cd(Cbar), info, ws, as_cd(cache));
}

size_t khx_redheffer_bwd_ws_bytes(int n, int batch) { return khx::star_bwd_ws_bytes(n, batch); }

int khx_redheffer_bwd(const double* SA, const double* SB, int n, int batch, const double* Sbar, double* SAbar,
                      double* SBbar, int* info, void* ws, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0 && ws && SA && SB && Sbar && SAbar && SBbar, "khx_redheffer_bwd: bad arguments");
  KHX_REQUIRE(batch <= 65535, "khx_redheffer_bwd: at most 65535 items per call");
  const long long s4 = 4ll * n * n;
  return khx::star_bwd(as_stream(stream), as_cd(SA), s4, as_cd(SB), s4, n, batch, as_cd(Sbar), as_cd(SAbar),
                       as_cd(SBbar), info, ws);
}

int khx_poynting_vec_adjoint(const double* sR, const double* sT, int N, const double* g, const double* kp,
                             const double* wl, const double* epsi_host, const double* epse_host, const double* wR,
                             const double* wT, int batch, double* sRbar, double* sTbar, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && batch >= 0 && epsi_host && epse_host, "khx_poynting_vec_adjoint: bad arguments");
  return khx::poynting_vec_adjoint(as_stream(stream), as_cd(sR), as_cd(sT), N, g, as_cd(kp), wl,
                                   mkc(epsi_host[0], epsi_host[1]), mkc(epse_host[0], epse_host[1]), wR, wT, batch,
                                   as_cd(sRbar), as_cd(sTbar));
}

int khx_eig_slots(int n) {
  KHX_ENTRY_BEGIN
  return khx::zeig_slots(n);
}

int khx_gemm_record(int on, int* shapes, int max_shapes) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(max_shapes >= 0, "khx_gemm_record: max_shapes < 0");
  return khx::gemm_record(on, shapes, max_shapes);
}

int khx_gemm_force(int cfg) {
  KHX_ENTRY_BEGIN
  return khx::gemm_force(cfg);
}

int khx_gemm_tune_set(int M, int N, int K, int opa, int opb, int batch, int cfg) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(M > 0 && N > 0 && K > 0 && opa >= 0 && opa <= 2 && opb >= 0 && opb <= 2 && batch > 0,
              "khx_gemm_tune_set: bad shape");
  return khx::gemm_tune_set(M, N, K, opa, opb, batch, cfg);
}

int khx_lu_factor(int n, double* A, int batch, int* ipiv, int* info, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && batch >= 0, "khx_lu_factor: bad sizes");
  return khx::zgetrf(as_stream(stream), n, khx::Mat{as_cd(A), n, (long long)n * n}, batch, ipiv, info);
}

int khx_lu_factor_solve(int n, int nrhs, double* A, double* B, int batch, int* ipiv, int* info, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && nrhs > 0 && batch >= 0, "khx_lu_factor_solve: bad sizes");
  const long long nn = (long long)n * n;
  return khx::zgetrf_solve(as_stream(stream), n, nrhs, khx::Mat{as_cd(A), n, nn},
                           khx::Mat{as_cd(B), nrhs, (long long)n * nrhs}, batch, ipiv, info);
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

int khx_zaxpby(double* D, int ldd, long long strideD, const double* a, const double* A, int lda, long long strideA,
               const double* b, const double* B, int ldb, long long strideB, int rows, int cols, int batch,
               void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(rows > 0 && cols > 0 && batch >= 0 && ldd >= cols && lda >= cols && (B == nullptr || ldb >= cols),
              "khx_zaxpby: bad sizes");
  const cd ca = mkc(a[0], a[1]), cb = mkc(b[0], b[1]);
  KHX_REQUIRE(B != nullptr || (cb.x == 0.0 && cb.y == 0.0), "khx_zaxpby: B is NULL but b != 0");
  return khx::zaxpby(as_stream(stream), khx::Mat{as_cd(D), ldd, strideD}, ca,
                     khx::Mat{const_cast<cd*>(as_cd(A)), lda, strideA}, cb,
                     khx::Mat{const_cast<cd*>(as_cd(B)), ldb, strideB}, rows, cols, batch);
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

int khx_mode_propagate_batch(const double* lam, const double* a, const double* zbar, int n, int nz, int batch,
                             double* M, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(n > 0 && nz >= 0 && batch >= 0, "khx_mode_propagate_batch: bad sizes");
  return khx::mode_propagate_batch(as_stream(stream), as_cd(lam), as_cd(a), zbar, n, nz, batch, as_cd(M));
}

int khx_longitudinal_batch(const double* kx, const double* ky, const double* F, int N, int nz, int batch, double* uz,
                           double* w, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && nz >= 0 && batch >= 0, "khx_longitudinal_batch: bad sizes");
  return khx::longitudinal_batch(as_stream(stream), as_cd(kx), as_cd(ky), as_cd(F), N, nz, batch, as_cd(uz),
                                 as_cd(w));
}

int khx_idft_phase_batch(const double* kx, const double* ky, const double* wl, const double* x, const double* y, int N,
                         int npts, int batch, double* E, void* stream) {
  KHX_ENTRY_BEGIN
  KHX_REQUIRE(N > 0 && npts >= 0 && batch >= 0, "khx_idft_phase_batch: bad sizes");
  return khx::idft_phase_batch(as_stream(stream), as_cd(kx), as_cd(ky), wl, x, y, N, npts, batch, as_cd(E));
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
