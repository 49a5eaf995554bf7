// Internal (non-exported) launchers shared by the khx translation units.
#pragma once
#include "khx_common.h"

namespace khx {

// Strided batch of row-major complex matrices: element (i, j) of item b is
// p[b * stride + i * ld + j].  stride 0 broadcasts one matrix to the batch.
struct Mat {
  cd* p;
  int ld;
  long long stride;
};
inline Mat mat(cd* p, int ld, long long stride) { return Mat{p, ld, stride}; }
inline Mat sub(Mat m, int r, int c) { return Mat{m.p + (long long)r * m.ld + c, m.ld, m.stride}; }

// C = alpha * A (M x K) * B (K x N) + beta * C
int zgemm(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch);
// C = alpha A^T (A stored K x M) * B + beta C
int zgemm_ta(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch);
// C = alpha A B + beta C with the product on the FP32 matrix cores (operands rounded to complex float)
int zgemm_f32(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch);
// C = alpha op(A) op(B) + beta C with op 0 = X, 1 = X^T, 2 = X^H
int zgemm_op(hipStream_t s, int opa, int opb, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C,
             int batch);
// GEMM tile selection (tuning table, forced choice, shape recording)
int gemm_tune_set(int M, int N, int K, int oa, int ob, int batch, int cfg);
int gemm_force(int cfg);
int gemm_record(int on, int* shapes, int max_shapes);
// Solve A X = B (A n x n destroyed, B n x m overwritten); info[b] = first zero pivot (1-based) or 0
int zgesv(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* info);

// LU factors with pivots (zgetrf layout) and the solve of their system
int zgetrf(hipStream_t s, int n, Mat A, int batch, int* ipiv, int* info);
int zgetrs(hipStream_t s, int n, int m, Mat A, const int* ipiv, Mat B, int batch);
int zgetrf_solve(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info);

// blocked LU (one workgroup per system, panel in LDS): keepL -> zgetrf layout
bool lu_blocked_fits(int n);
// multi-launch blocked LU (panel kernel + batched zgemm trailing updates)
bool lu_multi_fits(int n);
int zlu_multi(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info, bool keepL);
// zgesv with FP32-product triangular solves and one FP64 residual refinement step (m >= 64)
int zgesv_refined(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* info);
// LU path selection: 0 = unblocked, 1 = one-workgroup blocked, 2 = multi-launch blocked
int lu_path(int n);
// beam front end: paraxial beam fields on points, point DFT amplitudes
int beam_fields(hipStream_t s, int kind, const double* x, const double* y, const double* z, long long npts,
                double wl, double w0, double er, int p, int l, const double* geom, cd* E, cd* H);
int bloch_amplitudes(hipStream_t s, const cd* fields, const double* x, const double* y, int ns, int bz0, int bz1,
                     double kpx, double kpy, const double* gx, const double* gy, int nk, cd* out);
int points_dft(hipStream_t s, const cd* f, const double* x, const double* y, long long npts, const double* kx,
               const double* ky, int nk, int batch, double scale, cd* out);
// island layers: closed-form coefficients gathered into C (Fws: (2P-1)(2Q-1) complex)
// band-structure tool: S-matrix pencil split, Bloch eigenpairs, determinant
int scattering_splitlr(hipStream_t s, int ng, const cd* S, int batch, cd* Sl, cd* Sr);
size_t scattering_eig_ws_bytes(int ng, int batch);
int scattering_eig(hipStream_t s, int ng, const cd* S, int batch, cd sg, cd* w, cd* v, int* info, void* ws);
size_t scattering_det_ws_bytes(int ng, int batch);
int scattering_det(hipStream_t s, int ng, const cd* S, int batch, cd* det, int* info, void* ws);
int convmat_islands(hipStream_t s, const int* kinds, const double* params, const cd* eps, int nisl, cd eps_host,
                    const double* recip, double sigma, int P, int Q, cd* C, cd* Fws);
// B <- A^-1 B from zgetrf factors: interchanges, then blocked triangular solves
int zlu_solve_factored(hipStream_t s, int n, int m, Mat A, const int* ipiv, Mat B, int batch);
// B <- U^-1 B for the upper triangle U of A (its diagonal included)
int ztrsm_upper(hipStream_t s, int n, int m, Mat A, Mat B, int batch);
int zlu_blocked(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info, bool keepL);

// elementwise helpers (all batched, n_rows x n_cols views)
int zset_identity(hipStream_t s, Mat A, int rows, int cols, cd diag, int batch);   // A = diag * I
int zcopy(hipStream_t s, Mat dst, Mat src, int rows, int cols, int batch);
// D = a*A + b*B
int zaxpby(hipStream_t s, Mat D, cd a, Mat A, cd b, Mat B, int rows, int cols, int batch);
// D[i][j] = d[i] * A[i][j]   (d: batch x rows vector with stride dstride)
int zrow_scale(hipStream_t s, Mat D, const cd* d, long long dstride, Mat A, int rows, int cols, int batch);

// D = diag(dl) A diag(dr) (dl: batch x rows, dr: batch x cols; either may be null)
int zdiag_scale(hipStream_t s, Mat A, const cd* dl, const cd* dr, int rows, int cols, int batch, Mat D);

// adjoint helpers (adjoint.hip)
int conj_transpose(hipStream_t s, Mat A, int rows, int cols, Mat B, int batch);
int reduce_dot(hipStream_t s, Mat A, Mat B, int rows, int cols, int axis, int accumulate, cd* out,
               long long ostride, int batch);
int eig_adjoint_mix(hipStream_t s, cd* G, const cd* mu, const cd* mubar, int n, int batch, double eta_rel);
int lambda_adjoint(hipStream_t s, const cd* lam, const double* k0d, const cd* xbar, const cd* ilbar, int n,
                   int batch, cd* mubar);
int build_pq(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky, int N,
             int batch, cd* P, cd* Q);
int pq_adjoint(hipStream_t s, const cd* Pbar, const cd* Qbar, const cd* kx, const cd* ky, int N, int batch,
               cd* ICbar, cd* Cbar);

// field reconstruction for many sources at once (fields.hip)
int mode_propagate_batch(hipStream_t s, const cd* lam, const cd* a, const double* zbar, int n, int nz, int batch,
                         cd* M);
int longitudinal_batch(hipStream_t s, const cd* kx, const cd* ky, const cd* F, int N, int nz, int batch, cd* uz,
                       cd* w);
int idft_phase_batch(hipStream_t s, const cd* kx, const cd* ky, const double* wl, const double* x, const double* y,
                     int N, int npts, int batch, cd* E);

// eigen-decomposition: A destroyed; lam (n per item), W (n x n, unit columns).
// The workspace keeps Zt (the Schur vectors, transposed) and X (the unit upper
// triangular eigenvectors of the Schur form) after the call: W = Zt^T X diag(f),
// and rscale (optional, n per item) receives 1 / f.
size_t zeig_workspace_bytes(int n, int batch);
int zeig(hipStream_t s, int n, Mat A, cd* lam, long long lstride, Mat W, int batch, int* info, void* ws,
         unsigned long long* tphase = nullptr, cd* rscale = nullptr);
// matrices in flight at once on the current device (a full round of workgroups)
int zeig_slots(int n);
// zeig with its workspace in a static device array (no allocation, no ws argument)
int zeig_noalloc(hipStream_t s, int n, Mat A, cd* lam, long long lstride, Mat W, int batch, int* info);
size_t zeig_noalloc_limit();

// Raise a kernel's dynamic-LDS limit to `bytes` when it exceeds the 64 KiB
// default.  One cache per (device, kernel) shared by every caller; the
// attribute is only ever raised, so a later launch that needs less never
// lowers it under one that needs more.
int ensure_dynamic_lds(const void* fn, size_t bytes);

}  // namespace khx
