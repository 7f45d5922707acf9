/*
 * titan_hip.h -- C ABI of the MI355X (gfx950) implementation of TITAN's
 * chi(omega) path: H(k) assembly, batched G(k,z) = (z - H(k))^-1 on the
 * complex energy contour, the k- and energy-summed bare susceptibility
 * chi0(q,omega), and the RPA Dyson step chi = (I + chi0 U)^-1 chi0.
 *
 * Conventions (identical to TITAN's Fortran arrays, so a bind(C) interface
 * passes them without copies):
 *   - complex numbers are interleaved (re, im) doubles = complex(c_double_complex);
 *   - matrices are column-major: A(i,j) at A[i + j*lda];
 *   - Hamiltonian index (atom i, spin s, orbital mu) -> (2*i + s)*nOrb + mu
 *     (atom-major, spin up block then spin down block; reference 'ia' map,
 *     source/mod_System.F90:101-155);
 *   - chi index (sigma, i, mu, nu) -> 4*((i*nOrb + mu)*nOrb + nu) + sigma with
 *     sigma = 0..3 for (+, up, down, -) (reference 'sigmaimunu2i',
 *     source/mod_System.F90:158-223);
 *   - atoms and hop endpoints are 0-based.
 *
 * Status codes: 0 = success; > 0 = a pivot of (z - H) or (I + chi0 U) is
 * exactly zero (LAPACK 'info' semantics, as invers, source/mod_tools.F90:
 * 796-829); < 0 = error (see TH_E_* below).  The Fortran shim maps any
 * nonzero status to abortProgram (source/mod_mpi_pars.F90:59-72).
 *
 * Threading: a handle lives on the device that was current at th_create and
 * owns a HIP stream of its own; calls on one handle must not overlap, calls
 * on different handles may run concurrently from different host threads
 * (their kernels then overlap on the GPU).  The handle owns every device
 * buffer; the caller owns all host arrays, and no host pointer is retained
 * after a call returns.
 */
#ifndef TITAN_HIP_H
#define TITAN_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct th_handle th_handle;

#define TH_OK            0
#define TH_E_ARG        -1   /* illegal argument (size, null pointer, range) */
#define TH_E_NODEVICE   -2   /* no HIP device / library built without gfx950 code */
#define TH_E_ALLOC      -3   /* device allocation failed */
#define TH_E_HIP        -4   /* HIP runtime error (message via th_last_error) */
#define TH_E_STATE      -5   /* call order violated (e.g. chi0 before the k-mesh is set) */
#define TH_E_UNSUPPORTED -6  /* size outside what the kernels support (see th_limits) */

/* Precision modes of th_chi0 / th_chi_sweep. */
#define TH_PREC_FP64   0     /* G, chi0 contraction and sums in FP64 */
#define TH_PREC_MIXED  1     /* eigen-decomposition (or elimination) of H in FP64; the
                                Green function products and the chi0 contraction on the
                                fp16 matrix cores with two-term fp16 splits of FP32
                                values (FP32 accumulation, FP32 MFMA where H is not
                                Hermitian); FP64 contour sums across tiles and splits;
                                the block chi0[S,S] on the support S of U recomputed in
                                FP64 (see th_set_interaction); FP64 Dyson */

/* Creates a handle on the current HIP device and uploads the tight-binding
 * model: the local Hamiltonian h0 (d x d complex, d = 2*nOrb*nA) built by
 * hamilt_local (source/mod_hamiltonian.F90:35-87) and the hop table that
 * calchk sums (source/mod_hamiltonian.F90:183-243):
 *   H(k)[block jb[n], block ib[n]] += exp(i k.R_n) t_n   (both spins)
 * with R = 3*nhop cell vectors, t = nhop blocks of nOrb x nOrb complex
 * (column-major, rows = orbitals of atom jb[n]; TITAN's Neighbors(n)%t0i).
 * Replaces the set-up half of eintshechi (source/eintshechi.F90:36-44). */
int th_create(int nA, int nOrb, int nhop, const double *R, const int *jb, const int *ib,
              const double *t, const double *h0, th_handle **out);

/* Imaginary-axis (E,k) work list, the flattened E_k_imag_mesh of
 * genLocalEKMesh (source/adaptiveMesh.F90:76-161): point p has wave vector
 * k[3p..3p+2], energy y[p] (z = E_F + i(y + eta)) and weight w[p] = wght(E) *
 * w_k.  Meshes may differ per energy (adaptive meshes).  n >= 1. */
int th_set_imag_points(th_handle *h, long n, const double *k, const double *y, const double *w);

/* Real-axis k-mesh (realBZ%kp / realBZ%w, source/mod_BrillouinZone.F90:
 * 66-108), combined with the n_x real energies given to th_chi0. */
int th_set_real_kmesh(th_handle *h, long nk, const double *k, const double *w);

/* chi0(q, omega) for one frequency: the body of eintshechi(q, e)
 * (source/eintshechi.F90:2-178):
 *   chi0 = (1/2pi) [ sum_imag w (Gq G + (G Gq)^*)  +  sum_real ... ]
 * with Gq = G(k+q, E_F+omega+i(y+eta)), G = G(k, E_F+i(y+etap)) on the
 * imaginary axis and, when |omega| >= 1e-15, Gq = G(k+q, x+omega+i eta),
 * G = G(k, x+i etap) for x in x2 (weights p2, from generate_real_epoints,
 * source/EnergyIntegration.F90:99-122).  chi0_out is D x D complex,
 * D = 4*nOrb*nOrb*nA (chiorb_hf, source/mod_susceptibilities.F90:8). */
int th_chi0(th_handle *h, const double *q, double omega, double Ef, double eta, double etap,
            int n_x, const double *x2, const double *p2, int precision, double *chi0_out);

/* Declares the interaction matrix U (D x D complex, column-major, Umatorb of
 * build_identity_and_U_matrix, source/mod_susceptibilities.F90:106-159) that
 * the Dyson step will use; NULL clears it.  No counterpart in the reference:
 * the Dyson step chi = chi0 - chi0[:,S] U_S (I + chi0[S,S] U_S)^-1 chi0[S,:]
 * multiplies an error of the block chi0[S,S] on the support S of U by the
 * condition number of (I + chi0[S,S] U_S), which is large near the
 * Goldstone/FMR pole.  With U declared, TH_PREC_MIXED th_chi0 recomputes that
 * block in FP64 (where H is Hermitian and d is 18, 36 or 54: from the FP64
 * eigenpairs of H(k) with the energy sum taken per k-point first), so the
 * mixed chi keeps the accuracy of the mixed chi0 at any distance from the
 * pole.  th_chi_sweep in TH_PREC_MIXED declares its own U. */
int th_set_interaction(th_handle *h, const double *U);

/* RPA Dyson step chi = (I + chi0 U)^-1 chi0 in FP64 on the GPU
 * (source/calculate_chi.F90:84-89; U = Umatorb of
 * build_identity_and_U_matrix, source/mod_susceptibilities.F90:106-159).
 * Evaluated through the support of U (push-through identity), which gives
 * the same matrix.  cond_est (may be NULL) receives the 1-norm condition
 * number of the reduced matrix (I + chi0[S,S] U[S,S]) that carries all of
 * the near-pole ill-conditioning. */
int th_dyson(th_handle *h, const double *chi0, const double *U, double *chi_out, double *cond_est);

/* th_dyson without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  chi0_dev
 * and chi_dev are device pointers to D x D complex (column-major, as
 * th_dyson's host arrays); U is the host D x D interaction matrix.  Work
 * space (the support of U, U_S, the reduced matrix and the two rank-|S|
 * factors) lives in a fixed device array of 64 MiB per device, set up by
 * th_create: 2 |S| D + 2 |S|^2 complex values must fit (TH_E_UNSUPPORTED
 * otherwise).  Every call on one device shares that array, so calls that
 * overlap on one device (from different handles or threads) race; one call
 * at a time per device.  Same kernels and results as th_dyson; synchronises
 * the handle's stream before it returns. */
int th_dyson_noalloc(th_handle *h, const double *chi0_dev, const double *U, double *chi_dev, double *cond_est);

/* Full frequency sweep on the device (calculate_chi's omega loop,
 * source/calculate_chi.F90:66-107): for each omega[w] computes chi0 and chi
 * and their orbital traces
 *   schi[(sigma,i),(sigma',j)] = sum_{mu,nu} chi[(sigma,i,mu,mu),(sigma',j,nu,nu)]
 * (4nA x 4nA complex, index sigma*nA + i).  Real-axis nodes are generated on
 * the host per omega with parts3 x n3gl Gauss-Legendre points.  Any of the
 * outputs may be NULL.  chi_full_out, if given, receives the last omega's
 * full chi (D x D).  cond_out[w] as th_dyson. */
int th_chi_sweep(th_handle *h, const double *q, int n_omega, const double *omega, double Ef,
                 double eta, double etap, int parts3, int n3gl, const double *U, int precision,
                 double *schi0_out, double *schi_out, double *cond_out, double *chi_full_out);

/* Adjoint (reverse-mode derivative) of one frequency of the chi path.  No
 * counterpart in the reference, which has no derivatives; it differentiates
 * th_chi0 followed by th_dyson with respect to the model th_create uploaded
 * (hops t, local Hamiltonian h0) and to U.  Given the cotangent chi_hat of
 * the output (dL/dRe + i dL/dIm of a real loss L, D x D column-major), it
 * returns
 *   t_hat   nhop blocks nOrb x nOrb, same layout and hop order as th_create's t,
 *   h0_hat  d x d,
 *   U_hat   D x D (may be NULL),
 * in the same convention.  chi0 and chi are the forward results at this
 * omega (from th_chi0 and th_dyson with the same arguments).  With U == NULL
 * the Dyson step is skipped: chi_hat is then the cotangent of chi0 and chi is
 * not read.  The Green functions are recomputed in the backward, not stored
 * for it (th_tune "adjoint_bcache" = 1 reads the omega-independent
 * imaginary-axis ones from the forward's panel cache instead); 'precision'
 * selects the panel precision of the two adjoint contractions as in th_chi0
 * (TH_PREC_MIXED with "contract_f16": fp16 matrix cores on two-term splits;
 * the products G^H Ghat G^H, in the eigenbasis of H(k) when the eigenpairs
 * are kept, in FP32) while every sum is FP64. */
int th_chi_adjoint(th_handle *h, const double *q, double omega, double Ef, double eta, double etap,
                   int n_x, const double *x2, const double *p2, int precision,
                   const double *chi0, const double *U, const double *chi, const double *chi_hat,
                   double *t_hat, double *h0_hat, double *U_hat);

/* Releases every device buffer of the handle. */
/* Orbital traces of a susceptibility and the spin-orbit cross responses
 * (replaces the post-processing in calculate_chi.F90:91-185 that follows the
 * Dyson step for every frequency).
 *   chi     D x D complex128 column-major, or NULL for the chi of the last
 *           th_dyson / th_chi_sweep on this handle (no transfer)
 *   schi    out, 4nA x 4nA: schi[(s,i),(t,j)] = sum_{mu,nu}
 *           chi[(s,i,mu,mu),(t,j,nu,nu)], index s*nA + i (sigmai2i), s in
 *           (+, up, down, -); may be NULL
 *   angles  NULL, or per atom (theta, phi) in degrees of its magnetisation:
 *           schi is then rotated to the local frames, block (i,j) -> A_i
 *           block B_j with TITAN's rotation_matrices_chi at (180 - theta,
 *           180 + phi) (calculate_chi.F90:143-185; needs schi)
 *   lvec    NULL, or nOrb x nOrb x 3 complex128, lvec[mu + nOrb*(gamma +
 *           nOrb*s)] = <mu|L_s|gamma> (TITAN's Types%lvec)
 *   ls, sl, ll  out (need lvec), 3nA x 3nA each, index s*nA + i, s = x, y, z:
 *           <<L,S>>, <<S,L>>, <<L,L>> (calculate_chi.F90:109-140); may be NULL
 * Outputs are column-major complex128. */
int th_chi_traces(th_handle *h, const double *chi, const double *lvec, const double *angles, double *schi,
                  double *ls, double *sl, double *ll);

/* Local spin density <S^{mu nu}_ii>_k = <c+_{i mu} sigma_k c_{i nu}> of the
 * ground state by an imaginary-axis sweep over the handle's imaginary list
 * (replaces calcSmunu, mod_sumrule.F90:162-245, which the sum rule check
 * sumrule() of mod_sumrule.F90:10-109 consumes): S = sum_p w_p (G + G^H)/2 pi
 * over z = E_F + i(y_p + eta) plus 1/2 on the diagonal of the up-up and
 * down-down blocks.  S_out: nOrb x nOrb x 3 x nA complex128, S_out[mu +
 * nOrb*(nu + nOrb*(k + 3*i))] (TITAN's Smunuiivec), k = x, y, z. */
int th_spin_density(th_handle *h, double Ef, double eta, double *S_out);

/* Ground-state occupations and magnetisation per orbital by the same sweep
 * (replaces expectation_values_greenfunction, mod_expectation.F90:28-149,
 * without superconductivity): rho = n_up + n_down, mz = n_up - n_down with
 * n_s = 1/2 + sum_p w_p Re G[(i,s,mu),(i,s,mu)] / pi, mp = mx + i my =
 * sum_p w_p (G[(i,d,mu),(i,u,mu)] + conj G[(i,u,mu),(i,d,mu)]) / pi over
 * z = E_F + i(y_p + eta).  rho, mx, my, mz: nOrb x nA doubles, mp: nOrb x nA
 * complex128, element [mu + nOrb*i] (TITAN's rho(mu,i) etc.). */
int th_expectation_values(th_handle *h, double Ef, double eta, double *rho, double *mx, double *my, double *mz,
                          double *mp);

/* Parallel spin-current response to an electric field along ue[3] (replaces
 * eintshe(q,e), eintshe.F90:2-224, with dtdksub, dtdksub.F90:2-28; callers
 * calculate_all.F90:137, calculate_dc_limit.F90:134): over the handle's
 * imaginary list and, for omega != 0 with n_x > 0, the real k-mesh x the nodes
 * (x2, p2), Gq = G(k+q, z+omega) and G = G(k, z) as th_chi0 takes them,
 * contracted with the velocity vertex dtdk = sum_n i (ue.R_n) e^{i k.R_n} t_n
 * in the 16 spin-block terms of the reference, summed with the pair weights,
 * multiplied by prefactor (D x D complex128 column-major, e.g.
 * (I + chi_hf U)^-1; NULL = identity) and divided by 2 pi.
 * out: D x 4 complex128 column-major (TITAN's tchiorbiikl).  FP64, any d: up
 * to d = 54 one kernel per chunk of pairs holds a pair's G, Gq and products
 * in LDS; beyond (or with the tuning key "current_general") the products go
 * through device memory in three stages with the same sums. */
int th_current_response(th_handle *h, const double *q, double omega, double Ef, double eta, double etap,
                        int n_x, const double *x2, const double *p2, const double *ue, const double *prefactor,
                        double *out);

int th_destroy(th_handle *h);

/* Text of the last error on this thread (never NULL). */
const char *th_last_error(void);

/* Limits of this build: max atom-block size (2*nOrb), max d handled by the
 * LDS-resident inversion, max supported nA. */
int th_limits(int *max_nOrb, int *max_d, int *max_nA);

/* Device time of the most recent th_chi0 / th_chi_sweep / th_chi_adjoint call
 * split by stage (milliseconds, HIP events on the stream each stage runs on;
 * with "overlap" the inversion and contraction stages run concurrently, so
 * their sum can exceed the whole):
 * t[0] green-function kernels, t[1] contraction GEMM, t[2] reduction +
 * finalize, t[3] Dyson, t[4] whole call.  Also the number of kernel
 * launches of each kind in n[0..3]. */
int th_last_timing(th_handle *h, double *t, long *n);


/* Tuning knobs of the kernels (no counterpart in the reference; used by the
 * autotuner in titan_amd/autotune.py):
 *   "green_variant"  inversion kernel shape: 0 = 2-D register-blocked (default
 *                    tiles), 1 = 1-D ownership, 2 = one wave per matrix,
 *                    3 = two waves per matrix (2 and 3 where compiled for d)
 *   "spectral"       1 = where H(k) is Hermitian and d is 18, 36, 54, 72, 144
 *                    or 288, each distinct k is diagonalised once (Jacobi in
 *                    LDS for d <= 54, the batched dense solver above; at the
 *                    first frequency of every th_chi_sweep, th_chi0 keeps the
 *                    eigenpairs of earlier calls at the same q; for d <= 144
 *                    V is also kept as split fp16 planes for the fp16
 *                    kernels) and G(k,z) = V diag(1/(z-lambda)) V^H is formed
 *                    on the matrix cores for every z (FP64 MFMA; in
 *                    TH_PREC_MIXED fp16 MFMA on two-term splits with
 *                    "green_f16" = 1, the default, for d <= 144, else FP32
 *                    MFMA) (default); 0 = every G by Gauss-Jordan
 *                    elimination of z-H
 *   "spectral_chunk" 1 = the spectral kernel takes runs of up to 16 pairs of
 *                    one k-point per workgroup, loads V once per run and
 *                    writes the panels from the MFMA accumulators (default);
 *                    0 = one pair per workgroup (G staged in LDS)
 *   "pair_order"     1 = th_set_imag_points stores the list in runs of 16
 *                    pairs of one k-point: the r-th pair of a k-point goes to
 *                    block r/16 and the list is sorted by (block, k-point,
 *                    r); th_chi0 walks the real axis the same way (nodes in
 *                    blocks of 16, each block k-point by k-point), and the
 *                    Green function kernels store each batch's panel rows
 *                    with its runs transposed, the order in which the
 *                    contraction sums them (default); 0 = the given order.
 *                    Applies from the next
 *                    th_set_imag_points; chi0 changes only in the rounding
 *                    of its sum over pairs
 *   "overlap"        1 = th_chi0 pipelines the batches on two streams: the
 *                    eliminations of a batch run while the MFMA contraction
 *                    of the previous one runs (default; it applies when the
 *                    spectral path is not taken); 2 = the same, also for the
 *                    spectral Green functions; 0 = one stream.  Results are
 *                    bitwise the same for every value
 *   "contract_f16"   TH_PREC_MIXED: 1 = the contraction runs on the fp16 matrix
 *                    cores, each FP32 panel value split into two fp16 parts
 *                    (hi + lo, 22 bits) under one power-of-two scale per
 *                    operand from the resolvent bound of a Hermitian H, three
 *                    fp16 products per real product, FP32 accumulation
 *                    (default; FP32 MFMA where H is not Hermitian); 0 = FP32
 *                    MFMA
 *   "presplit"       TH_PREC_MIXED with contract_f16, where both sides take the
 *                    half-precision spectral kernel: 1 = that kernel writes
 *                    the (hi, lo) fp16 words itself and the contraction copies
 *                    them (default; the same values as 0); 0 = FP32 panels
 *                    split in the contraction
 *   "contract_tiles" contraction of presplit panels: 1 = eight waves per
 *                    112 x 112 tile, two per SIMD, each holding a 2-D block of
 *                    6 or 7 of the 49 16 x 16 output blocks, panels prefetched
 *                    two stages ahead (default); 0 = seven waves, one 16-row
 *                    strip each.  Results are bitwise the same for both
 *   "green_f16"      TH_PREC_MIXED spectral path: 1 = V diag(d) V^H on the fp16
 *                    matrix cores with two-term splits (default; d <= 144: the
 *                    run kernel for d <= 54, the column-tiled kernel with the
 *                    diagonal on its B side for d = 72 and 144); 0 = FP32 MFMA
 *   "contract_3m"    1 = the contraction forms each complex product from three
 *                    real MFMA products (Ar Br, Ai Bi, (Ar+Ai)(Br+Bi)) instead
 *                    of four; 0 = four (default)
 *   "graph"          1 = th_chi_sweep captures one frequency step (chi0, Dyson,
 *                    traces) as a hipGraph at its first frequency with and
 *                    without real axis and replays it for the others, omega
 *                    read from device memory; per-stage times are then not
 *                    split (only t[4] is set).  0 = plain launches (default)
 *   "splits"         K-splits of the contraction (workgroups per tile)
 *   "cache_mb"       cap (MB) of this handle's omega-independent caches, the
 *                    imaginary-axis panel cache and the eigenpair sets
 *                    (0 = none: the panel cache may take 70 % and the
 *                    eigenpairs 30 % of free HBM); a cache over the cap is
 *                    not kept (panels recomputed per frequency, Green
 *                    functions by elimination), so many handles can share HBM
 *   "drop_caches"    1: forgets the omega-independent imaginary-axis panel
 *                    cache and the eigenpairs, so that the next call computes
 *                    them again into the allocations it keeps (timing from a
 *                    cold start without the one-time allocations)
 *   "current_general" th_current_response: 1 = the three-stage form through
 *                    device memory for every d; 0 = the one-kernel form where a
 *                    pair fits the LDS (d <= 54, default).  The sums are the
 *                    same up to the order of the two imaginary-axis terms
 *   "refine"         TH_PREC_MIXED with a declared U (th_set_interaction):
 *                    1 = chi0[S,S] on the support of U recomputed in FP64
 *                    (default); 0 = left as the mixed contraction gives it;
 *                    2 = as 1, and TH_PREC_FP64 too takes chi0[S,S] from the
 *                    eigenpair form (a second FP64 evaluation of the block
 *                    that the Dyson step amplifies, to measure how far two
 *                    FP64 summations of chi disagree at a pole)
 *   "batch_mb"       panel memory per batch in MB of FP64-sized panels (pairs
 *                    per launch); a budget of at least one full wave of the
 *                    fp16 spectral kernel's workgroups (d = 18, 36, 54) is
 *                    rounded to the nearest whole number of waves
 *   "adjoint_bcache" 0 = th_chi_adjoint recomputes every Green function
 *                    (default); 1 = it reads the omega-independent
 *                    imaginary-axis B panels from the forward's panel cache
 *                    when that holds them (same E_F, eta', precision)
 * th_get_tuning reads the current values. */
int th_tune(th_handle *h, const char *key, long value);
int th_get_tuning(th_handle *h, long *green_variant, long *splits, long *batch_mb);

/* Kernel-level test hooks (no counterpart in the reference; used by the
 * GPU unit tests to check one kernel at a time).
 * thx_green_panel: runs the green-function kernel in 'mode' (0 imag A,
 *   1 imag B, 2 real A, 3 real B) for pairs p0..p0+np-1 and returns the
 *   panel [np][nA*nA][2][Mpad] as doubles, Mpad = ceil((2 nOrb)^2/112)*112.
 *   Takes the path th_chi0 takes under the handle's tuning (spectral and
 *   pair_order: pairs in the stored order, real axis grouped when set).
 * thx_contract: uploads panels A, B of that layout, runs the contraction and
 *   returns the accumulator [nA*nA][2][Mpad][Mpad] (re block, im block).
 * thx_last_split: 1 when the last chi0 had the Green function kernels write
 *   split half-precision panels, else 0.
 * thx_last_refined: 1 when the last chi0 recomputed chi0[S,S] in FP64. */
int thx_green_panel(th_handle *h, int mode, long p0, long np, const double *q, double omega, double Ef,
                    double eta, double etap, int prec, int n_x, const double *x2, const double *p2, double *out);
int thx_contract(th_handle *h, const double *A, const double *B, long np, int prec, double *acc_out);
int thx_last_split(th_handle *h);
int thx_last_refined(th_handle *h);
/* thx_last_spectral: 1 when the last chi0 took its Green functions from the
 * eigenpairs of H(k) (spectral kernels), 0 when from elimination. */
int thx_last_spectral(th_handle *h);
/* thx_eigenpairs: the eigenpairs of H(k + q) at the handle's unique k-points
 * (the imaginary list's, then the real mesh's new ones) as the spectral
 * kernels use them: L_out n_kuniq x d, V_out n_kuniq x d x d complex, V[r][m]
 * = component r of eigenvector m.  Computes them where they are not valid. */
int thx_eigenpairs(th_handle *h, const double *q, double *L_out, double *V_out);

#ifdef __cplusplus
}
#endif
#endif
