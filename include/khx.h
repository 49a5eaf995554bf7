/*
 * khx -- C-ABI of the MI355X RCWA solve path (libkhx.so, gfx950).
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer.  Complex arrays are
 *     interleaved (re, im) double pairs: numpy/torch complex128 layout.
 *   - Matrices are row-major and batch-major: item b of a batch of n x n
 *     matrices starts at b*n*n complex elements.  S-matrices are stored
 *     [batch][2][2][n][n] (blocks S11, S12, S21, S22).
 *   - n = 2*N where N = P*Q is the number of harmonics; harmonic h = q*P + p
 *     with p (the x index) fastest; a field vector is [x part; y part].
 *   - All calls are asynchronous on `stream` (a hipStream_t; NULL = default
 *     stream), allocate nothing and never synchronise.  Scratch memory is
 *     passed in `ws` with the size given by the matching *_ws_bytes query.
 *   - Return value: 0 on success, -1 invalid argument, -3 HIP error; the
 *     message is in khx_last_error() (thread-local).  Numerical status per
 *     batch item is written to `info` (device int array) where noted.
 *
 * The reference (Kaeryv/Khepri) has no FFI: its boundary is the Python
 * Crystal/Layer API.  Each entry point below names the reference function it
 * replaces; the Python package khepri_amd binds them with ctypes.
 */
#ifndef KHX_H
#define KHX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* khx_last_error(void);
/* ABI version of the library (incremented on signature changes). */
int khx_abi_version(void);

/* k vectors of the expansion (khepri/expansion.py:43-50 _compute_k_vector,
 * x/y rows): kx[b][h] = (kp[b][0] + g[0][h]) / k0_b, ky likewise, with
 * k0_b = 2 pi / wl[b].
 *   g   : real (2, N)            kp : complex (batch, 2)
 *   wl  : real (batch)           kx, ky : complex (batch, N)               */
int khx_kvectors(const double* g, int N, const double* kp, const double* wl, int batch, double* kx,
                 double* ky, void* stream);

/* Convolution (Toeplitz) matrix of `count` pixmaps
 * (khepri/tools.py:33-56 convolution_matrix):
 *   C[q*P+p][q'*P+p'] = fftshift(fft2(eps))[nx/2 + p-p'][ny/2 + q-q'] / (nx*ny)
 * eps: (count, nx, ny) real (is_complex = 0) or complex (is_complex = 1),
 * indexed eps[x][y].  C: complex (count, N, N).  Fails (-1) when the pixmap
 * is too small for the harmonics, where the reference raises IndexError.   */
/* Beam front end (khepri/beams.py).  khx_beam_fields: E, H (each 3 x npts
 * complex, component-major) of an x-polarised paraxial beam, kind 0 =
 * Gaussian (w0 = beam waist), 1 = Laguerre-Gauss LG_pl, evaluated at the
 * device points (x, y, z) taken to the beam frame; geom (device, 21 doubles)
 * = rotation matrix M (row-major), M^-1, M^-1 o (o = beam origin); replaces
 * shifted_rotated_fields with _paraxial_gaussian_field_fn /
 * _paraxial_laguerre_gaussian_field_fn (khepri/beams.py:50-160).
 * khx_points_dft: out[b][k] = scale sum_p f[b][p] exp(-i (kx[k] x[b][p] +
 * ky[k] y[b][p])) for batch point sets of npts points (slow_dft,
 * khepri/fourier.py:94-104, as used by amplitudes_from_fields,
 * khepri/beams.py:164-190).                                                 */
/* khx_bloch_amplitudes: the whole of amplitudes_from_fields
 * (khepri/beams.py:164-190) on the device: fields (R, C, 2, 3) complex and
 * coordinates x, y (R, C) sampled on bz0 x bz1 square cells of ns x ns points
 * (R = ns bz1, C = ns bz0); the Bloch phase of (kpx, kpy) is removed at each
 * field point and out (4 x nk complex: Ex, Ey, Hx, Hy) holds the sums over
 * every cell of the field cells' DFTs at the harmonics (gx, gy), with the
 * reference's pairing of field cells and coordinate cells, / (bz0 bz1 ns).  */
int khx_bloch_amplitudes(const double* fields, const double* x, const double* y, int ns, int bz0, int bz1,
                         double kpx, double kpy, const double* gx, const double* gy, int nk, double* out,
                         void* stream);
int khx_beam_fields(int kind, const double* x, const double* y, const double* z, long long npts, double wl,
                    double w0, double er, int p, int l, const double* geom, double* E, double* H, void* stream);
int khx_points_dft(const double* f, const double* x, const double* y, long long npts, const double* kx,
                   const double* ky, int nk, int batch, double scale, double* out, void* stream);

/* Band-structure tool (khepri/eigentricks.py:5-55).  S (batch, 4ng, 4ng)
 * complex on the device in the (+, -) block layout: p = rows/columns 0:2ng,
 * m = 2ng:4ng, S[p,p] forward transmission, S[m,p] reflection from the left,
 * S[p,m] reflection from the right, S[m,m] backward transmission.
 * khx_scattering_splitlr: Sl = [[Spp, 0], [Smp, -I]], Sr = [[I, -Spm],
 * [0, -Smm]] (scattering_splitlr, khepri/eigentricks.py:5-22).
 * khx_scattering_eig: Bloch factors w (batch, 4ng) and unit-norm modes v
 * (batch, 4ng, 4ng, column j for w[j]) of the pencil Sl v = w Sr v
 * (scattering_eigenvalues, khepri/eigentricks.py:29-40, which calls scipy's
 * QZ eig(Sl, Sr)), computed as the eigenpairs of the Cayley transform
 * (Sl - s Sr)^-1 (Sl + s Sr); s = (sigma_re, sigma_im) on the unit circle must
 * not be a Bloch factor.  info[b]: 0 ok, > 0 eigensolver did not converge,
 * < 0 s is a Bloch factor of item b (call again with another s),
 * KHX_INFO_NAN_INPUT S[b] holds a NaN (the reference returns None); w and v
 * of items with info < 0 or KHX_INFO_NAN_INPUT are NaN.
 * khx_scattering_det: det[b] = det(Spp - Spm Smm^-1 Smp) det(Smm)
 * (scattering_det, khepri/eigentricks.py:24-28), formed as sign * exp(sum
 * log|u_ii|) like numpy's det; info[b] = k+1 when Smm is singular at pivot k
 * (numpy's inv raises LinAlgError).                                         */
#define KHX_INFO_NAN_INPUT 0x40000000
int khx_scattering_splitlr(int ng, const double* S, int batch, double* Sl, double* Sr, void* stream);
size_t khx_scattering_eig_ws_bytes(int ng, int batch);
int khx_scattering_eig(int ng, const double* S, int batch, double sigma_re, double sigma_im, double* w, double* v,
                       int* info, void* ws, void* stream);
size_t khx_scattering_det_ws_bytes(int ng, int batch);
int khx_scattering_det(int ng, const double* S, int batch, double* det, int* info, void* ws, void* stream);

/* Convolution matrix of an island layer (replaces the Formulation.ANALYTICAL
 * branch of Layer.solve, khepri/layer.py:161-166, with the island transforms
 * of khepri/fourier.py:7-58 and combine_fourier_masks :145-159).  Islands on
 * the device: kinds[i] 0 = disc (params c0, c1, r, unused), 1 = rectangle
 * (ll0, ll1, ur0, ur1), 2 = uniform (no term); eps complex per island;
 * eps_host complex (host pointer, 2 doubles); recip = b1x, b1y, b2x, b2y
 * (host, 4 doubles); sigma = unit-cell area.  C (P Q x P Q complex) on the
 * device; ws = khx_convmat_ws_bytes(P, Q, 1) bytes of device memory.      */
int khx_convmat_islands(const int* kinds, const double* params, const double* eps, int nisl,
                        const double* eps_host, const double* recip, double sigma, int P, int Q, double* C, void* ws,
                        void* stream);
size_t khx_convmat_ws_bytes(int P, int Q, int count);
int khx_convmat(const double* eps, int is_complex, int nx, int ny, int P, int Q, int count, double* C,
                void* ws, void* stream);

/* Batched complex LU solve A X = B (numpy.linalg.solve / inv as used in
 * khepri/layer.py:158 IC = inv(C) and throughout khepri/alternative.py).
 * A (batch, n, n) is overwritten by its factors, B (batch, n, nrhs) by X.
 * info[b] = k+1 when the k-th pivot is exactly zero (numpy: LinAlgError). */
int khx_zgesv(int n, int nrhs, double* A, double* B, int batch, int* info, void* stream);

/* Strided-batched complex GEMM C = alpha A B + beta C (numpy matmul as used
 * throughout khepri/alternative.py).  Item b of A is the M x K row-major
 * matrix at A + b*strideA with row stride lda (complex elements); a batch
 * stride of 0 shares one operand.  alpha, beta: host complex scalars.      */
int khx_zgemm(int M, int N, int K, const double* alpha, const double* A, int lda, long long strideA,
              const double* B, int ldb, long long strideB, const double* beta, double* C, int ldc,
              long long strideC, int batch, void* stream);

/* Batched eigen-decomposition of general complex matrices
 * (numpy.linalg.eig / LAPACK zgeev as called at khepri/alternative.py:172).
 * A (batch, n, n) is destroyed; lam (batch, n) eigenvalues; W (batch, n, n)
 * eigenvectors in columns, unit 2-norm, largest component real.
 * info[b] > 0: QR iteration did not converge.                             */
size_t khx_eig_ws_bytes(int n, int batch);
int khx_eig(int n, double* A, int batch, double* lam, double* W, int* info, void* ws, void* stream);
/* khx_eig that allocates nothing and takes no workspace: the same kernels,
 * with their scratch (khx_eig_ws_bytes(n, batch) - 256 bytes: the Schur
 * vectors and triangular eigenvectors, 2 n^2 x 16 bytes per matrix) in a
 * static device array of khx_eig_noalloc_bytes() = 64 MiB that each GPU the
 * library runs on holds until the process ends.  No counterpart in the
 * reference: it is for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Fails (-1) when the batch does not fit the array (n = 162: up to 39
 * matrices).  Every call on one GPU shares that GPU's array: calls on one
 * stream are ordered, calls that overlap in time on different streams of
 * one GPU race.  Nothing in the array needs to start at zero.              */
int khx_eig_noalloc(int n, double* A, int batch, double* lam, double* W, int* info, void* stream);
size_t khx_eig_noalloc_bytes(void);
/* khx_eig that also writes, per matrix, 10 words to tphase (device,
 * batch x 10): wall-clock stamps (100 MHz) at start / after the Hessenberg
 * reduction / after the Schur iteration / after the eigenvectors, the
 * number of QR iterations (low 32 bits; the chain passes among them in the
 * high 32 bits), and (H in global memory) the ticks spent in the
 * deflation scan and shift, the in-block chase, the rest of the blocked
 * sweeps, the rotation log and its flushes -- measurement support only.     */
int khx_eig_timed(int n, double* A, int batch, double* lam, double* W, int* info, void* ws,
                  unsigned long long* tphase, void* stream);

/* Patterned layer: modes + S-matrix for a batch of (kx, ky, k0*depth) with
 * the convolution matrix C and IC = C^-1 shared by the batch (per_item = 0:
 * C, IC are N x N) or one per item (per_item = 1: (batch, N, N)).
 * Replaces khepri/layer.py:157-160 (Layer.solve, FFT formulation):
 *   khepri/alternative.py:158-178 solve_structured_layer -> (W, V, lam)
 *   khepri/alternative.py:181-195 build_scatmat          -> S
 * k0d[b] = k0_b * depth.  S: (batch, 2, 2, n, n).  W, V (batch, n, n) and
 * lam (batch, n) are written when non-NULL (needed for fields).
 * info[b] > 0: eigen-solver did not converge.                              */
size_t khx_patterned_layer_ws_bytes(int n, int batch);
int khx_patterned_layer(const double* C, const double* IC, int per_item, const double* kx, const double* ky,
                        int N, int batch, const double* k0d, double* S, double* W, double* V, double* lam,
                        int* info, void* ws, void* stream);
/* The same, also keeping the four solves' results for the backward twin
 * (khx_patterned_layer_bwd_cached): cache (batch, 9, n, n) = W^-1, V^-1 V0,
 * [G1 | G2] = A^-1 [X | B] (n x 2n), [S11 | S12] (n x 2n), A^-1, T^-1, V^-1.  Same reference
 * code as khx_patterned_layer; for gradients, where the forward pass's
 * intermediates are needed again.                                          */
int khx_patterned_layer_cached(const double* C, const double* IC, int per_item, const double* kx, const double* ky,
                               int N, int batch, const double* k0d, double* S, double* W, double* V, double* lam,
                               int* info, double* cache, void* ws, void* stream);

/* Uniform layer of permittivity eps (khepri/layer.py:166-168:
 * khepri/alternative.py:130-156 solve_uniform_layer + build_scatmat) and the
 * half spaces (khepri/layer.py:169-178: khepri/alternative.py:32-56
 * scattering_reflection, side = 1; :59-82 scattering_transmission, side = 2).
 * Every harmonic decouples into a 2x2 system; S, W, V are written dense.
 * eps: one complex scalar (2 doubles, host memory).  k0d ignored for half
 * spaces (may be NULL).                                                     */
int khx_uniform_layer(const double* kx, const double* ky, int N, int batch, const double* eps_host,
                      const double* k0d, double* S, double* W, double* V, double* lam, void* stream);
int khx_halfspace(int side, const double* kx, const double* ky, int N, int batch, const double* eps_host,
                  double* S, double* W, double* V, double* lam, void* stream);

/* Redheffer star product S = SA * SB (khepri/alternative.py:19-30
 * redheffer_product), batched; a_bcast / b_bcast share one operand.  S must
 * not alias SA or SB.  info[b] = first zero pivot of D or F (0 = none).    */
size_t khx_redheffer_ws_bytes(int n, int batch);
int khx_redheffer(const double* SA, int a_bcast, const double* SB, int b_bcast, int n, int batch, double* S,
                  int* info, void* ws, void* stream);

/* Reflected and transmitted power (khepri/crystal.py:363-396
 * poynting_flux_end, with khepri/alternative.py:101-128 incident and
 * :235-245 poynting_fluxes).  Stot (batch, 2, 2, n, n); g real (2, N);
 * kp complex (batch, 2); wl (batch); te, tm complex (batch);
 * epsi, epse: complex scalars in host memory.  R, T (batch); Rord, Tord
 * (batch, N) per diffraction order when non-NULL.                          */
size_t khx_poynting_ws_bytes(int N, int batch);
int khx_poynting(const double* Stot, int N, const double* g, const double* kp, const double* wl,
                 const double* te, const double* tm, const double* epsi_host, const double* epse_host, int batch,
                 double* R, double* T, double* Rord, double* Tord, void* ws, void* stream);

/* The same split in its two halves, for the R/T-only chain:
 * khx_incident writes the incident vector c1p (batch, n) -- zero except the
 * x / y parts of the central harmonic (khepri/alternative.py:101-128);
 * khx_poynting_vec takes the output amplitudes sR = Stot11 c1p and
 * sT = Stot21 c1p, each (batch, n) (khepri/alternative.py:235-245).       */
int khx_incident(int N, const double* kp, const double* wl, const double* te, const double* tm,
                 const double* epsi_host, int batch, double* c, void* stream);
int khx_poynting_vec(const double* sR, const double* sT, int N, const double* g, const double* kp,
                     const double* wl, const double* epsi_host, const double* epse_host, int batch, double* R,
                     double* T, double* Rord, double* Tord, void* stream);

/* LU factorisation with stored pivots (LAPACK zgetrf layout, 0-based ipiv)
 * and the matching solve (zgetrs) -- numpy.linalg.solve split in two, for a
 * factor reused across several solves.  A, LU (batch, n, n); ipiv (batch, n)
 * device ints; B (batch, n, nrhs) overwritten by X.                        */
int khx_lu_factor(int n, double* A, int batch, int* ipiv, int* info, void* stream);
int khx_lu_solve(int n, int nrhs, const double* LU, const int* ipiv, double* B, int batch, void* stream);
/* Both at once: A -> its factors (as khx_lu_factor), B -> A^-1 B.          */
int khx_lu_factor_solve(int n, int nrhs, double* A, double* B, int batch, int* ipiv, int* info, void* stream);

/* Elementwise helpers of the S-matrix algebra: out = diag(dl) A diag(dr)
 * (strided batches like khx_zgemm; dl (batch, rows), dr (batch, cols),
 * either may be NULL; out may alias A)
 * and out = I - A (out may alias A).                                       */
int khx_diag_scale(const double* A, int lda, long long strideA, const double* dl, const double* dr, int rows,
                   int cols, int batch, double* out, int ldo, long long strideO, void* stream);
int khx_identity_minus(const double* A, int n, int batch, double* out, void* stream);
/* D = a A + b B on strided batches of (rows x cols) views (B may be NULL
 * when b = 0; D may alias A or B).  a, b: host complex scalars.            */
int khx_zaxpby(double* D, int ldd, long long strideD, const double* a, const double* A, int lda, long long strideA,
               const double* b, const double* B, int ldb, long long strideB, int rows, int cols, int batch,
               void* stream);

/* Propagation factors X (batch, n) of a uniform layer of the gap medium
 * (eps = 1), whose S-matrix is exactly [[0, diag X], [diag X, 0]]
 * (khepri/alternative.py:130-156 + 181-195 with W = W0, V = V0).          */
int khx_propagation(const double* kx, const double* ky, int N, int batch, const double* k0d, double* X,
                    void* stream);

/* Extended (twisted-bilayer) basis: scatter the base-layer matrices solved at
 * kp + g_m (m = 0..Nb-1, one per harmonic of the other lattice) into the
 * 2Nb^2 x 2Nb^2 joint space (khepri/extension.py:10-63 _joint_subspace /
 * joint_subspace).  in: (items, Nb, blocks, 2Nb, 2Nb) -- blocks = 4 for
 * S-matrices, 1 for W / V; out: (items, blocks, 2Nb^2, 2Nb^2), fully written.
 * kind 0: base is the right-hand expansion, 1: the left-hand one.
 * The _diag variant scatters eigenvalue vectors (items, Nb, 2Nb) ->
 * (items, 2Nb^2) (khepri/extension.py:110).                                */
int khx_joint_subspace(const double* in, int Nb, int blocks, int kind, int items, double* out, void* stream);
int khx_joint_subspace_diag(const double* in, int Nb, int kind, int items, double* out, void* stream);

/* Field reconstruction (khepri/crystal.py:234-343, khepri/fields.py,
 * khepri/fourier.py:136-142).
 *   khx_eigenbasis:     R = [[W, W], [-V, V]]  (batch, 2n, 2n)   fields.py:46-51
 *   khx_mode_propagate: M[i][j] = L_i(zbar_j) a_i, L = [e^{lam zbar}, e^{-lam zbar}]
 *                       with |L| clamped to 1e14; lam (n), a (2n), M (2n, nz)
 *                                                                fields.py:29-31, 53-62
 *   khx_longitudinal:   F = [sx; sy; ux; uy] (4N, nz) -> uz = -i (kx sy - ky sx)
 *                       and w = kx uy - ky ux (sz = -i IC w)     fields.py:68-76
 *   khx_idft_phase:     E[p][g] = exp(i k0 (Kx_g x_p + Ky_g y_p)), (npts, N);
 *                       Kx, Ky normalised (as from khx_kvectors)  fourier.py:136-142 */
int khx_eigenbasis(const double* W, const double* V, int n, int batch, double* R, void* stream);
int khx_mode_propagate(const double* lam, const double* a, const double* zbar, int n, int nz, double* M,
                       void* stream);
int khx_longitudinal(const double* kx, const double* ky, const double* F, int N, int nz, double* uz, double* w,
                     void* stream);
int khx_idft_phase(const double* kx, const double* ky, double k0, const double* x, const double* y, int N, int npts,
                   double* E, void* stream);
/* The same three for `batch` sources at once (the Brillouin-zone loop of
 * khepri's examples/bzi/twisted_phc.py:126-147 runs fields_volume once per
 * Bloch vector; here every Bloch vector / wavelength is one item):
 *   khx_mode_propagate_batch: lam (batch, n), a (batch, 2n), zbar (nz) shared
 *                             -> M (batch, 2n, nz)             fields.py:29-31, 53-62
 *   khx_longitudinal_batch:   kx, ky (batch, N), F (batch, 4N, nz)
 *                             -> uz, w (batch, N, nz)          fields.py:68-76
 *   khx_idft_phase_batch:     kx, ky (batch, N) normalised, wl (batch) device,
 *                             x, y (npts) shared -> E (batch, npts, N) with
 *                             k0 = 2 pi / wl[b]                fourier.py:136-142 */
int khx_mode_propagate_batch(const double* lam, const double* a, const double* zbar, int n, int nz, int batch,
                             double* M, void* stream);
int khx_longitudinal_batch(const double* kx, const double* ky, const double* F, int N, int nz, int batch, double* uz,
                           double* w, void* stream);
int khx_idft_phase_batch(const double* kx, const double* ky, const double* wl, const double* x, const double* y, int N,
                         int npts, int batch, double* E, void* stream);

/* ---------------------------------------------------------------- gradient
 * Reverse-mode (adjoint) building blocks for d(R, T)/d(pixel eps).  The
 * reference has no gradient; these are the adjoints of the forward entry
 * points above (khepri/tools.py:33-56 convolution_matrix,
 * khepri/alternative.py:158-195 structured layer + build_scatmat,
 * :19-30 redheffer_product, :235-245 poynting_fluxes).  Convention:
 * Zbar = dL/dRe Z + i dL/dIm Z for a real objective L.
 *
 * khx_zgemm_op: C = alpha op(A) op(B) + beta C, op 0 = X, 1 = X^T, 2 = X^H
 *   (a transposed operand is stored with its other dimension first).
 * khx_conj_transpose: B = A^H for (rows x cols) items.
 * khx_reduce_dot: out[b][i] (+)= sum_j A_ij conj(B_ij) (axis 1) or
 *   out[b][j] (+)= sum_i A_ij conj(B_ij) (axis 0); out item stride ostride.
 * khx_eig_adjoint_mix: G (batch, n, n) holds W^H Wbar on entry and
 *   G_ij E_ij / (|E_ij|^2 + eta^2), E_ij = mu_j - mu_i, G_ii = mubar_i on exit
 *   (eta = eta_rel max|mu|: Lorentzian broadening of degenerate pairs).
 * khx_lambda_adjoint: mubar from lambar contributions of x = exp(-lam k0d)
 *   (xbar) and 1/lam (ilbar), lam = sqrt(mu); either may be NULL.
 * khx_build_pq: dense P and Q (batch, n, n) of a patterned layer.
 * khx_pq_adjoint: ICbar, Cbar (batch, N, N) from Pbar, Qbar.
 * khx_convmat_adjoint: pixel adjoint epsbar (batch, nx, ny) (real, or
 *   interleaved complex when is_complex) from Cbar (batch, N, N).
 * khx_poynting_vec_adjoint: sRbar, sTbar (batch, n) for the objective
 *   sum_b wR[b] R[b] + wT[b] T[b] (wR or wT may be NULL).                   */
int khx_zgemm_op(int opa, int opb, int M, int N, int K, const double* alpha, const double* A, int lda,
                 long long strideA, const double* B, int ldb, long long strideB, const double* beta, double* C,
                 int ldc, long long strideC, int batch, void* stream);
int khx_conj_transpose(const double* A, int lda, long long strideA, int rows, int cols, double* B, int ldb,
                       long long strideB, int batch, void* stream);
int khx_reduce_dot(const double* A, int lda, long long strideA, const double* B, int ldb, long long strideB,
                   int rows, int cols, int axis, int accumulate, double* out, long long ostride, int batch,
                   void* stream);
int khx_eig_adjoint_mix(double* G, const double* mu, const double* mubar, int n, int batch, double eta_rel,
                        void* stream);
int khx_lambda_adjoint(const double* lam, const double* k0d, const double* xbar, const double* ilbar, int n,
                       int batch, double* mubar, void* stream);
int khx_build_pq(const double* C, const double* IC, int per_item, const double* kx, const double* ky, int N,
                 int batch, double* P, double* Q, void* stream);
int khx_pq_adjoint(const double* Pbar, const double* Qbar, const double* kx, const double* ky, int N, int batch,
                   double* ICbar, double* Cbar, void* stream);
size_t khx_convmat_adjoint_ws_bytes(int P, int Q, int batch);
int khx_convmat_adjoint(const double* Cbar, int P, int Q, int nx, int ny, int batch, int is_complex,
                        double* epsbar, void* ws, void* stream);
/* Backward twins of khx_patterned_layer and khx_redheffer.
 * khx_patterned_layer_bwd: Cbar (batch, N, N) = adjoint of the convolution
 *   matrix (IC = C^-1 is differentiated too) given Sbar (batch, 2, 2, n, n),
 *   the modes W, V, lam written by khx_patterned_layer for the same inputs
 *   and eta_rel, the Lorentzian broadening of the eigenvector adjoint
 *   (1e-12 is a good default).  The S-matrix intermediates are recomputed.
 * khx_redheffer_bwd: SAbar, SBbar (batch, 2, 2, n, n) for S = SA * SB given
 *   Sbar (no broadcast operands).  info as in the forward calls.           */
size_t khx_patterned_layer_bwd_ws_bytes(int n, int batch);
int khx_patterned_layer_bwd(const double* C, const double* IC, int per_item, const double* kx, const double* ky,
                            int N, int batch, const double* k0d, const double* W, const double* V,
                            const double* lam, const double* Sbar, double eta_rel, double* Cbar, int* info, void* ws,
                            void* stream);
/* khx_patterned_layer_bwd with the forward intermediates taken from the cache
 * written by khx_patterned_layer_cached for the same inputs (no re-solve). */
int khx_patterned_layer_bwd_cached(const double* C, const double* IC, int per_item, const double* kx,
                                   const double* ky, int N, int batch, const double* k0d, const double* W,
                                   const double* V, const double* lam, const double* cache, const double* Sbar,
                                   double eta_rel, double* Cbar, int* info, void* ws, void* stream);
size_t khx_redheffer_bwd_ws_bytes(int n, int batch);
int khx_redheffer_bwd(const double* SA, const double* SB, int n, int batch, const double* Sbar, double* SAbar,
                      double* SBbar, int* info, void* ws, void* stream);
int khx_poynting_vec_adjoint(const double* sR, const double* sT, int N, const double* g, const double* kp,
                             const double* wl, const double* epsi_host, const double* epse_host, const double* wR,
                             const double* wT, int batch, double* sRbar, double* sTbar, void* stream);

/* Number of matrices khx_eig keeps in flight at once on the current device
 * (resident workgroups per CU x CUs; no counterpart in the reference).  A
 * batch that is a multiple of it leaves no partly filled last round; callers
 * that stream a large batch in chunks size the chunks by it.  Returns the
 * count (>= 1) or a negative status.                                        */
int khx_eig_slots(int n);

/* GEMM tile autotuning (no counterpart in the reference).  Tile shapes:
 * 0 = 64x64, 1 = 32x32, 2 = 64x32, 3 = 32x64 (M x N per workgroup, FP64 FMA),
 * 4..7 = the same tiles on the FP64 matrix cores (MFMA).
 * khx_gemm_record(1, NULL, 0) starts recording the distinct GEMM shapes the
 * library runs; khx_gemm_record(0, shapes, max) stops and copies up to max
 * records of 6 ints (M, N, K, opA, opB, batch) into shapes, returning the
 * count.  khx_gemm_force(cfg) uses tile cfg for every GEMM (-1: off);
 * khx_gemm_tune_set installs the choice for one shape class (batch is
 * bucketed: < 16, < 256, >= 256; cfg -1 removes it).                       */
int khx_gemm_record(int on, int* shapes, int max_shapes);
int khx_gemm_force(int cfg);
int khx_gemm_tune_set(int M, int N, int K, int opa, int opb, int batch, int cfg);

/* Measurement support (no counterpart in the reference).
 * khx_prof_enable(1) clears the records and brackets every launch of the
 * instrumented launches (zeig: the eigensolver's kernels of one call;
 * zgesv_kernel, zgemm_kernel) with
 * a HIP event pair on its stream; khx_prof_collect sums the durations of one
 * kernel (synchronising on its events).  khx_copy is a plain device copy of
 * `bytes` (multiple of 16) used to calibrate HBM byte counters.            */
int khx_prof_enable(int on);
int khx_prof_collect(const char* kernel, double* total_ms, long long* launches);
int khx_copy(const void* src, void* dst, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* KHX_H */
