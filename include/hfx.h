/* libhfx — MI355X (gfx950) Hartree-Fock Fock-build / direct-minimisation kernels.
 *
 * Plain C ABI: caller-owned device pointers (complex data interleaved re/im,
 * complex128 for HFX_C128 plans and complex64 for HFX_C64 plans; all real
 * reductions, weights and moments are float64), sizes as ints, a hipStream_t
 * passed as void*.  Every function returns 0 on success or a negative status
 * (HFX_EINVAL bad argument, HFX_EHIP runtime error, HFX_EUNSUPPORTED shape);
 * hfx_last_error() returns the thread-local message of the last failure.
 * Array layout is the reference's: (nk1, nk2, nb, nb) row-major, centred
 * meshgrid order; per-k vectors (nk1, nk2, nb); weights (nk1, nk2).
 *
 * Each entry point names the reference interface (src/jax_hf/...) it replaces.
 */
#ifndef HFX_H
#define HFX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFX_C128 0
#define HFX_C64 1

typedef struct hfx_plan hfx_plan;

/* ABI version of this header (bumped on any signature change). */
int hfx_abi_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated). */
int hfx_last_error(char* buf, size_t n);

/* Plan for an (nk1, nk2) mesh with nb bands in the given storage precision.
 * Uploads the FFT twiddle tables and the full / upper-triangle channel lists.
 * Replaces the shape-keyed jit cache of _solve_jitted (solver.py:663-683). */
int hfx_plan_create(int nk1, int nk2, int nb, int dtype, hfx_plan** out);
int hfx_plan_destroy(hfx_plan* plan);

/* Exchange-kernel variants for the compile-time square meshes (12 .. 96):
 * HFX_EXCH_V1 runs one Stockham stage per prime-power radix; HFX_EXCH_V2_A/B/C
 * run two composite-radix stages per axis with workgroups of 1024/768/512
 * (large meshes) threads; HFX_EXCH_V3_A/B/C (slot planes only, else the V2 of
 * the same size) fuse the VR multiply with the last forward and the first
 * inverse column stage.  HFX_EXCH_AUTO is the built-in per-mesh default.
 * Other meshes always use the runtime-radix kernel. */
#define HFX_EXCH_AUTO 0
#define HFX_EXCH_V1 1
#define HFX_EXCH_V2_A 2
#define HFX_EXCH_V2_B 3
#define HFX_EXCH_V2_C 4
#define HFX_EXCH_V3_A 5
#define HFX_EXCH_V3_B 6
#define HFX_EXCH_V3_C 7

/* Selects the exchange-kernel variant used by every later call on this plan.
 * No counterpart in the reference (XLA chooses its own FFT plan). */
int hfx_plan_set_exchange_variant(hfx_plan* plan, int variant);
int hfx_plan_get_exchange_variant(const hfx_plan* plan);

/* Times every exchange variant (reps launches each, all nb*nb channels, on the
 * caller's buffers; Sigma is overwritten) and keeps the fastest on the plan.
 * planar: P / Sigma are slot planes and VR channel-major (hfx_selfenergy_planar).
 * times_ms (optional, 8 doubles) receives the mean time per launch of variants
 * 0 .. 7 (0 = AUTO's choice is not timed and is reported as 0).  Synchronises
 * the stream.  No counterpart in the reference (XLA autotunes its own FFTs). */
int hfx_exchange_autotune(hfx_plan* plan, const void* P, const void* VR, int vr_full, const int* vr_map, int planar,
                          void* Sigma, int reps, double* times_ms, int* best, void* stream);

/* Bytes of caller-provided float64 scratch needed by the energy reductions
 * (calls sharing one scratch must not overlap). */
size_t hfx_energy_scratch_bytes(const hfx_plan* plan);

/* Exchange self-energy on a channel list (selfenergy_fft, utils.py:66-173 and
 * the block-spec subsets utils.py:176-311):
 *   Sigma_c = -IFFT2[FFT2(P_c - refP_c) * VR_c],  VR phase-absorbed (problem.py:88-98).
 * chans == NULL selects the plan's full list (herm = 0) or upper triangle
 * (herm = 1, lower triangle written as the conjugate mirror, utils.py:140-161).
 * ifftshift = 1 applies the final ifftshift of selfenergy_fft (utils.py:135-136).
 * vr_full = 1 for VR of shape (nk1, nk2, nb, nb), 0 for (nk1, nk2, 1, 1).
 * rho (optional, complex128 nb*nb) receives sum_k w_k (P - refP)_c. */
int hfx_selfenergy(const hfx_plan* plan, const void* P, const void* refP, const void* VR, int vr_full,
                   const int* chans, int nchan, int herm, int ifftshift, const double* w2d, void* Sigma,
                   void* rho, void* stream);

/* VR = fftn(w[..,None,None] * Vq, axes=(0,1)) (problem.py:88), times the
 * ifftshift phase exp(2 pi i (n0 s0/N0 + n1 s1/N1)) when phase = 1
 * (problem.py:90-98).  Vq complex in the plan precision, (nk1, nk2, nb, nb);
 * a scalar (nk1, nk2, 1, 1) Vq uses a plan created with nb = 1. */
int hfx_prepare_vr(const hfx_plan* plan, const void* Vq, int vr_full, const double* w2d, int phase,
                   void* VR, void* stream);

/* Adjoint of the exchange map (backward of build_fock / hf_energy, no reference
 * counterpart: jax autodiff differentiates fock.py:51-64): for Sigma = A[P],
 * A_c = IFFT2 diag(-VR_c) FFT2, writes A^H[G] = -IFFT2[conj(VR_c) FFT2(G_c)]
 * on all nb^2 channels -- the cross-correlation, recomputed by the same fused kernel. */
int hfx_selfenergy_adjoint(const hfx_plan* plan, const void* G, const void* VR, int vr_full, void* out,
                           void* stream);

/* out_c = FFT2(w * X_c) over all nb^2 channels (w float64 (nk1, nk2)); the forward
 * transforms of the VR gradient dE/dVR = -conj(FFT2 P) FFT2(G) / N. */
int hfx_fft2_weighted(const hfx_plan* plan, const void* X, const double* w, void* out, void* stream);

/* Weighted channel moments rho = sum_k w_k (P - refP) (fock.py:66-70, 79). */
int hfx_density_moments(const hfx_plan* plan, const void* P, const void* refP, const double* w2d,
                        void* rho, void* stream);

/* k-independent Hartree diagonal H_a = sum_b HH_ab Re rho_bb (fock.py:66-74) and
 * contact self-energy sum_t g_t [O_i tr(O_j rho) - O_i rho O_j] (fock.py:76-92).
 * HH, Hdiag or (g, Oi, Oj, sigc) may be NULL to skip that term. */
int hfx_kterms(const hfx_plan* plan, const void* rho, const double* HH, int ntc, const double* g,
               const void* Oi, const void* Oj, double* Hdiag, void* sigc, void* stream);

/* Fock epilogue of build_fock (fock.py:94) fused with hf_energy (fock.py:109-111):
 *   F = herm(h + Sigma + sigc + diag(Hdiag)),  Sigma_out = Sigma + sigc,
 *   E_out = sum_k w_k Re Tr[(h + (Sigma_out + diag(Hdiag))/2) P] + e_offset.
 * Sigma, Hdiag, sigc, Sigma_out, P/E_out optional (NULL).  Sigma_out must not
 * alias Sigma when sigc is given. scratch: hfx_energy_scratch_bytes(). */
int hfx_fock_epilogue(const hfx_plan* plan, const void* h, const void* Sigma, const double* Hdiag,
                      const void* sigc, const void* P, const double* w2d, void* F, void* Sigma_out,
                      double* scratch, double* E_out, double e_offset, void* stream);

/* ---- planar (channel-major) Fock build --------------------------------------
 * The same build as hfx_selfenergy + hfx_fock_epilogue (fock.py:76-111), with
 * the exchange operands staged as slot planes Xt[s][t][k] = X[s][k][chans[t]]
 * so that every exchange load and store is coalesced.  chans NULL selects the
 * upper triangle (herm) or all nb^2 channels.  Replaces the same reference
 * calls as the (k, i, j)-layout entry points above. */
/* X (nsolve, nk1, nk2, nb, nb) [- refX (nk1, nk2, nb, nb)] -> Xt (nsolve, nchan, nk1, nk2).
 * Also makes the channel-major VR (herm 0, chans NULL) that hfx_selfenergy_planar reads. */
int hfx_planar_pack(const hfx_plan* plan, int nsolve, const void* X, const void* refX, const int* chans, int nchan,
                    int herm, void* Xt, void* stream);
/* Sigma_t = exchange of the slot planes Pt (nsolve, nchan, nk1, nk2); VRc the
 * channel-major VR (nb^2, nk1, nk2) when vr_full, else (nk1, nk2); rho (nb, nb)
 * weighted moments of Pt (nsolve == 1 only, mirrored with herm).  vr_map
 * (optional, device int32 nb^2, vr_full only): channel c reads VR plane
 * vr_map[c] -- a layer-resolved V that depends only on the layer pair has a
 * few distinct planes, which are then read from HBM once per build instead of
 * once per channel (no counterpart in the reference, whose VR is dense). */
int hfx_selfenergy_planar(const hfx_plan* plan, int nsolve, const void* Pt, const void* VRc, int vr_full,
                          const int* vr_map, const int* chans, int nchan, int herm, int ifftshift,
                          const double* w2d, void* Sigt, void* rho, void* stream);
/* hfx_fock_epilogue reading Sigma from slot planes (mirrored with herm;
 * channels outside a given chans list are zero). */
int hfx_fock_epilogue_planar(const hfx_plan* plan, const void* h, const void* Sigt, const int* chans, int nchan,
                             int herm, const double* Hdiag, const void* sigc, const void* P, const double* w2d,
                             void* F, void* Sigma_out, double* scratch, double* E_out, double e_offset,
                             void* stream);

/* hf_energy (fock.py:100-111) on given Sigma / H; H broadcast (1,1,nb,nb) when
 * H_bcast = 1, full (nk1,nk2,nb,nb) when 0, absent when NULL. */
int hfx_hf_energy(const hfx_plan* plan, const void* P, const void* h, const void* Sigma, const void* H,
                  int H_bcast, const double* w2d, double* scratch, double* E_out, void* stream);

/* ---- utilities ---------------------------------------------------------- */
#define HFX_KIND_F64 0
#define HFX_KIND_F32 1
#define HFX_KIND_C128 2
#define HFX_KIND_C64 3

/* Elementwise precision conversion of n values between storage kinds
 * (the mixed-precision plumbing of the precision form; real -> complex sets
 * im = 0, complex -> real keeps the real part). */
int hfx_convert(const void* src, int src_kind, void* dst, int dst_kind, size_t n, void* stream);

/* out = a X + b Y elementwise (a = ar + i ai, b = br + i bi; Y may be NULL);
 * the density mixing / DIIS extrapolation / Steihaug vector updates of
 * reference_scf.py:275-339 and solver.py:947-989. out may alias X or Y. */
int hfx_lincomb(size_t n, int kind, double ar, double ai, const void* X, double br, double bi, const void* Y,
                void* out, void* stream);

/* out = sum_v coef[v] X[v] over nmat nb x nb matrices (1 <= nin <= 8, real
 * coefficients), Hermitised per matrix when herm: the DIIS extrapolation
 * F = herm(sum_i c_i F_i) of reference_scf.py:298-314 in one pass. out may
 * alias an input only without herm. */
int hfx_lincomb_multi(size_t nmat, int nb, int kind, int nin, const void* const* X, const double* coef, int herm,
                      void* out, void* stream);

/* out = 1 / (1 + exp((x - mu) / (T + 1e-12))), mu = *mu or 0 (fermidirac, utils.py:51-53). */
int hfx_fermi(const double* x, const double* mu, double T, size_t n, double* out, void* stream);

/* out[k,c] = a conj(X[k,c]) Y[k,c] (sum_ch = 0) or out[k] = a sum_c conj(X[k,c]) Y[k,c]
 * (sum_ch = 1, nch channels per k): the pointwise products of the VR gradient. */
int hfx_cmul(int dtype, size_t nk, int nch, const void* X, const void* Y, double ar, double ai, int sum_ch,
             void* out, void* stream);

/* out[k, c] = alpha w[k] X[k, c] (w may be NULL): the k-weighting of adjoint seeds. */
int hfx_scale_k(size_t nk, int nch, int kind, const void* X, const double* w, double alpha, void* out,
                void* stream);
/* out[k] = alpha (*alpha_dev) w[k] op(X[k]) for nk complex nb x nb matrices
 * (kind 2 = complex128, 3 = complex64): op = conjugate transpose when dagger
 * (out must not alias X then); alpha_dev (a device scalar, e.g. the incoming
 * gradient of the energy) and w optional.  The w P^dagger terms of the
 * build_fock / hf_energy adjoint (jax autodiff of fock.py:100-111). */
int hfx_scale_k_ex(size_t nk, int nb, int kind, const void* X, const double* w, double alpha, const double* alpha_dev,
                   int dagger, void* out, void* stream);

/* Y = 0.5 (X + X^dagger) over nmat nb x nb matrices (fock.py:13-15, 95-96). */
int hfx_hermitize(const void* X, void* Y, size_t nmat, int nb, int dtype, void* stream);
/* hfx_hermitize from X of x_dtype into Y of y_dtype (HFX_C128 / HFX_C64): the
 * Hermitised start density of solver.py:380-394 in the loop's precision, one
 * pass.  Y must not alias X. */
int hfx_hermitize_convert(const void* X, int x_dtype, void* Y, int y_dtype, size_t nmat, int nb, void* stream);

/* out[0] = max|X|, out[1] = max|X| over channels with mask[c] != 0; the
 * off-block check of the block-spec exchange (utils.py:263-266, 282, 295).
 * scratch: >= 1024 doubles. */
int hfx_absmax_masked(const void* X, size_t nmat, int nb, int dtype, const unsigned char* mask,
                      double* scratch, double* out, void* stream);

/* occupation_entropy (fock.py:114-118) of p (nk, nb) with weights w (nk);
 * with E (device scalar) non-NULL writes free_energy E - max(T,1e-14) S
 * (fock.py:121-129) instead.  real_kind: HFX_KIND_F64 or HFX_KIND_F32. */
int hfx_entropy(const void* p, int real_kind, const double* w, size_t nk, int nb, const double* E, double T,
                double* scratch, double* out, void* stream);

/* ---- per-k batched kernels (nb <= 32: one wavefront per k-point up to nb = 16,
 * a 256-thread workgroup per k-point above) ----------------------------- */
/* Device scalar state of the solver loops: float64[HFX_ST_LEN]. */
#define HFX_ST_MU 0
#define HFX_ST_BETA 1
#define HFX_ST_LAM 2
#define HFX_ST_GRAD 3
#define HFX_ST_TAU0 4
#define HFX_ST_TAU 5
#define HFX_ST_TAUF 6
#define HFX_ST_ACC 7
#define HFX_ST_OMEGA0 8
#define HFX_ST_DNORM 9
#define HFX_ST_OMEGA 10
#define HFX_ST_E 11
#define HFX_ST_ITER 12   /* iteration counter advanced by the HFX_OP_BETA step */
#define HFX_ST_MULO 13   /* mu bracket and done flag of hfx_solve_mu_steps */
#define HFX_ST_MUHI 14
#define HFX_ST_MUDONE 15
#define HFX_ST_FROZEN 16 /* batched solves: a converged solve's kernels skip their updates */
#define HFX_ST_LEN 32

/* Batched Hermitian eigensolver (jnp.linalg.eigh call sites, solver.py:165,
 * 389, 623, 842; reference_scf.py:89): cyclic round-robin Jacobi, eigenvalues
 * ascending in lam (nk, nb), eigenvectors as columns of V.  herm_in = 1
 * hermitises A first.  dtype: HFX_C128 / HFX_C64. */
int hfx_eigh_batched(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, int herm_in,
                     void* stream);
/* hfx_eigh_batched with the Jacobi rotations stopped once every off-diagonal
 * |A_pq| <= rel_tol ||A||_F (rel_tol below machine epsilon: the full
 * precision of hfx_eigh_batched).  For eigendecompositions inside an
 * iteration whose own residual test decides convergence (the DIIS polish
 * updates); rel_tol <= 1e-3. */
int hfx_eigh_batched_tol(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, int herm_in,
                         double rel_tol, void* stream);
/* The eigendecomposition of A = U^dagger F U started from the basis U, as
 * the reference's warm eigh of the rotated Fock matrix followed by the change
 * of basis (solver.py:622-625, reference_scf.py:87-96): (lam, W) = eigh(A)
 * to rel_tol (as hfx_eigh_batched_tol) and V = U W, the rotations
 * accumulated onto U's rows (no W is formed); V may be U (in place).  A must
 * not alias U or V. */
int hfx_eigh_batched_basis(int dtype, int nb, size_t nk, const void* A, double* lam, const void* U, void* V,
                           int herm_in, double rel_tol, void* stream);

/* P = herm(Q diag(p) Q^dagger) and part[k] = sum |P[k] - Pref[k]|^2: the
 * density of the orbitals (Q, p) of F[Pref] and its per-k density residual
 * (reference_scf.py:287-294, the Roothaan step's |P_new - P|). */
int hfx_density_diff(int dtype, int nb, size_t nk, const void* Q, const double* p, const void* Pref, void* P,
                     double* part, void* stream);
/* P = herm(Q diag(p) Q^dagger) (solver.py:42-44 + 462, utils.py:473-474);
 * when p is NULL, p = sigmoid((*mu - eps)/T) is formed (and stored to pout). */
int hfx_density_from_Qp(int dtype, int nb, size_t nk, const void* Q, const double* p, const double* eps,
                        const double* mu, double T, double* pout, void* P, void* stream);

/* CG iteration pieces (solver.py:451-600); part arrays are (nval, nk) float64. */
int hfx_cg_grad(int dtype, int nb, size_t nk, const void* Q, const void* F, const double* p, void* Ft,
                double* eps, void* G, double* part, void* stream);
int hfx_cg_precond(int dtype, int nb, size_t nk, const void* G, const void* Gprev, const double* eps,
                   const double* p, const double* gp_prev, const double* st, double T, void* H, double* gp,
                   double* hp, double* part, int seg_len, void* stream);
int hfx_cg_direction(int dtype, int nb, size_t nk, const void* H, void* D, const double* hp, double* dp,
                     const void* Ft, const double* st, void* V, double* lam_d, void* Fe, double* part,
                     void* stream);
int hfx_cg_trial(int dtype, int nb, size_t nk, const void* V, const double* lam_d, const void* Fe,
                 const double* p, const double* dp, const double* eps, const double* st, int omega0,
                 double* part, void* stream);
/* Q <- Q V diag(c(tau)) V^dagger (solver.py:586-590); tau = st[HFX_ST_TAUF]
 * unless use_tau; eps_new (spectral diag, Fe = V^dagger Ft V) and/or
 * Ft_out = U^dagger Fe U (Fe = Ft, solver.py:1005-1009) optional. */
int hfx_cg_retract(int dtype, int nb, size_t nk, void* Q, const void* V, const double* lam_d, const void* Fe,
                   const double* st, double tau, int use_tau, double* eps_new, void* Ft_out, void* stream);
/* The CG direction / trial / retraction without the spectral setup: the
 * Cayley factor U(tau) = (I + (tau/2) d)^{-1} (I - (tau/2) d), d = skew(D), is
 * formed per trial by Gauss-Jordan elimination (the LU form of the same
 * unitary, solver.py:140-197 and linalg.py Cayley retraction), so no eigh runs
 * inside the iteration.  direction: D <- H + beta D, dp <- hp + beta dp,
 * part = |D|^2; trial: as hfx_cg_trial from D and Ft; retract: Q <- Q U(st
 * [HFX_ST_TAUF]), eps_new = diag(U^dagger Ft U)  (solver.py:528-592). */
int hfx_cg_direction_gj(int dtype, int nb, size_t nk, const void* H, void* D, const double* hp, double* dp,
                        const double* st, double* part, int seg_len, void* stream);
int hfx_cg_trial_gj(int dtype, int nb, size_t nk, const void* D, const void* Ft, const double* p,
                    const double* dp, const double* eps, const double* st, int omega0, double* part,
                    void* U_keep, double* eps_keep, int seg_len, void* stream);
/* U_keep / eps_keep (optional, both or neither): the last evaluated trial's U
 * and diag(U^dagger Ft U); the retraction reuses them when a trial was
 * accepted (st[HFX_ST_ACC] set, so tau_final is that trial's tau). */
int hfx_cg_retract_gj(int dtype, int nb, size_t nk, void* Q, const void* D, const void* Ft, const double* st,
                      double* eps_new, const void* U_keep, const double* eps_keep, int seg_len, void* stream);
/* seg_len (hfx_cg_precond and the _gj kernels): 0 for one solve; otherwise the
 * nk k-points are consecutive solves of seg_len k-points each, and k-point k
 * reads the state vector st + HFX_ST_LEN * (k / seg_len) (batched scans). */

/* Q <- Q W with (eps, W) = eigh(Q^dagger F Q)  (solver.py:622-625). */
int hfx_rotate_eigh(int dtype, int nb, size_t nk, void* Q, const void* F, double* eps, void* stream);
/* out = opA(A) opB(B - Bsub) [opC(C)] per k, optionally hermitised; ops bit0/1/2
 * = conjugate-transpose A/B/C (the einsum products of solver.py:42-49, 918-924). */
int hfx_matmul_batched(int dtype, int nb, size_t nk, const void* A, const void* B, const void* C, int ops,
                       const void* Bsub, int herm_out, void* out, void* stream);

/* Ft = U^dagger herm(sum_v coef[v] X[v]) U per k (Hermitised; 1 <= nin <= 8):
 * the DIIS-extrapolated Fock matrix (reference_scf.py:298-314) expressed in the
 * basis U, ready for an eigh that starts from the previous orbitals. */
int hfx_extrap_rotate(int dtype, int nb, size_t nk, int nin, const void* const* X, const double* coef, const void* U,
                      void* Ft, void* stream);
/* hfx_extrap_rotate over a stack of solves of seg_len k-points each, with
 * per-solve coefficients on the device: coef_dev[s * 8 + v] for solve s. */
int hfx_extrap_rotate_batched(int dtype, int nb, size_t nk, int seg_len, int nin, const void* const* X,
                              const double* coef_dev, const void* U, void* Ft, void* stream);
/* U <- U W (in place) and P = herm(U diag(occ) U^dagger) per k with the new U:
 * eigenvectors back from the rotated basis and the density of solver.py:42-44. */
int hfx_rotate_density(int dtype, int nb, size_t nk, void* U, const void* W, const double* occ, void* P,
                       void* stream);

/* Symmetry averaging of per-k matrices, one group average of make_project_fn
 * (symmetry.py:46-150: unitary conjugation sums, the same-k / flipped-k
 * group, the spatial average over k-index maps, time reversal):
 *   out[k] = mask[k] ? weight sum_g U_g op_g(A[src_g(k)]) U_g^dagger : A[k]
 * for ng <= 64 elements: Ug (ng, nb, nb) in the storage precision, src
 * (ng, nk) int32 linear source k-points (the mod / flip negations and index
 * maps as tables; entries outside [0, nk) read A[k]), op_g = complex
 * conjugation when bit g of conj_bits is set, mask (nk) bytes or NULL.  out
 * must not alias A. */
int hfx_symmetry_average(int dtype, int nb, size_t nk, int ng, const void* A, const void* Ug, const int* src,
                         unsigned long long conj_bits, const unsigned char* mask, double weight, void* out,
                         void* stream);

/* Periodic linear resampling of a centred (n0, n1, ...) k-grid array onto an
 * (nk, nk, ...) grid (utils.py:477-535, resample_kgrid with method "linear"):
 * axis 0 then axis 1, source coordinate (j - nk/2) * (n_old/nk) + n_old/2
 * rounded in the array's real precision, neighbours taken modulo n_old; an
 * axis already of length nk is copied.  m = elements per k-point (nb * nb for
 * matrices), kind 0 = float64, 1 = float32, 2 = complex128, 3 = complex64.
 * out must not alias X.  The coarse-to-fine seed of solve_continuation
 * (continuation.py:48-51). */
int hfx_resample_kgrid(int kind, int n0, int n1, size_t m, int nk, const void* X, void* out, void* stream);

/* Superlattice Fock build on the extended momentum grid (superlattice.py:75-377:
 * build_extended_layout / make_superlattice_fock_fn / make_superlattice_build_fock_fn).
 * P, refP (NULL = 0), h, Sigma, F: (nkx, nky, D, D), D = n_G * dim_orb, in the
 * plan's precision (dtype HFX_C128 / HFX_C64).  Exchange (planes != NULL): the
 * dG channel blocks of P - refP are scattered into planes (n_delta, dim_orb^2,
 * Nx, Ny) at tile g_off[i] of plane pair_to_delta[i][j] for every (i, j) in
 * the CSR lists delta_i / delta_j (planes must be zero-filled by the caller
 * before the first call: a build writes only the pair blocks) -- convolved by
 * the (Nx, Ny, dim_orb) plan with the single plane V_lag_fft into sigma_planes
 * (-ifft2(fft2(X) V_lag_fft)), and gathered back through pair_to_delta (n_G,
 * n_G) and g_off (n_G, 2) tile offsets.  Hartree (hartree != 0): the
 * k-independent shift degeneracy * sum_c HH_orb[A, B, a, c] rho_c(dG(A, B))
 * on the orbital diagonal, with rho_c(d) the w2d-weighted k sum of P - refP
 * over the pairs of d (CSR delta_start / delta_i / delta_j); delta_scratch
 * n_delta * dim_orb complex128.  Sigma = exchange + Hartree, F = herm(h + Sigma).
 * Limits: n_G^2, n_delta * dim_orb^2 and nkx * nky <= 65535. */
int hfx_superlattice_fock(int dtype, int nkx, int nky, int n_G, int dim_orb, int n_delta, int Nx, int Ny,
                          const void* P, const void* refP, const int* pair_to_delta, const int* g_off, const hfx_plan* plan, const void* V_lag_fft,
                          void* planes, void* sigma_planes, int hartree, const double* w2d,
                          const int* delta_start, const int* delta_i, const int* delta_j, const double* HH_orb,
                          double degeneracy, void* delta_scratch, const void* h, void* Sigma, void* F,
                          void* stream);

/* Newton-Schulz re-orthonormalisation of nk nb x nb bases in place, steps in
 * [1, 4] times U <- U (3 I - U^dagger U) / 2 (each step squares the deviation
 * from unitarity; the CG orbitals handed to the FP64 polish, solver.py:42-44). */
int hfx_orthonormalize(int dtype, int nb, size_t nk, void* U, int steps, void* stream);
/* hfx_orthonormalize of complex64 orbitals Q32 into complex128 U (converted on
 * the way in): the FP32 CG orbitals that start the FP64 polish basis. */
int hfx_orthonormalize_from(int nb, size_t nk, const void* Q32, void* U, int steps, void* stream);

/* SCF residuals (reference_scf.py:291-294): comm = F D - D F (stored if comm),
 * part[0][k] = sum |comm|^2, part[1][k] = sum |Pn - D|^2 (Pn may be NULL). */
int hfx_scf_resid(int dtype, int nb, size_t nk, const void* F, const void* D, const void* Pn, void* comm,
                  double* part, void* stream);
/* hfx_scf_resid plus, in the same pass, the new row of the DIIS B matrix
 * (reference_scf.py:110-111): part[2 + v][k] = Re sum conj(comm) R_v for the
 * nr <= 8 stored residuals Rs (a HOST array of device pointers; Rs[self] is
 * the buffer comm is written to, its value is part[0][k]; self = -1: none).
 * comm is required.  hist_dtype: storage of comm and Rs -- dtype, or
 * HFX_C64 for a complex128 build at nb = 16 (the residual norms are formed in
 * FP64 before the store; the B row reads the FP32 history). */
int hfx_scf_resid_dots(int dtype, int nb, size_t nk, const void* F, const void* D, const void* Pn, void* comm,
                       const void* const* Rs, int nr, int self, int hist_dtype, double* part, void* stream);
/* part[v][k] = Re sum conj(X_v) Y_v (mode 0, the inner products of solver.py:205-208 and the
 * DIIS B matrix reference_scf.py:110-111) or Re tr(X_v Y_v) (mode 1, reference_scf.py:319);
 * Xs, Ys are HOST arrays of nv <= 8 device pointers. */
int hfx_batched_dots(int dtype, int nb, size_t nk, const void* const* Xs, const void* const* Ys, int nv, int mode,
                     double* part, void* stream);
/* part[v][k] = sum_b x_v[k,b] y_v[k,b] for per-k real vectors (solver.py:211-213). */
int hfx_vec_dots(const double* const* xs, const double* const* ys, int nv, size_t nk, int nb, double* part,
                 void* stream);
/* x[i] -= scale * (*s), s a device scalar (the projection of solver.py:224-233). */
int hfx_vec_shift(double* x, size_t n, const double* s, double scale, void* stream);
/* x[i] -= scale * sum_k w[k] part[k] (part: per-k sums, nk of them), the sum
 * formed inside the launch: the projection of solver.py:224-233 in one launch. */
int hfx_vec_shift_sum(double* x, size_t n, const double* part, const double* w, size_t nk, double scale,
                      void* stream);
/* Up to four hfx_lincomb_dev updates of arrays of one shape in one launch:
 * op j: out1[j] = a_j' X1[j] + b_j' Y1[j] (n1 elements of kind1; Y1[j] NULL:
 * zero) and the same on the second arrays (n2 of kind2), a_j' = a[j] *
 * dev_scalars[a_idx[j]] when a_idx[j] >= 0, else a[j] (b alike).  The ops are
 * applied in order at each element, so op j may read an array that a later
 * op writes: the four updates of one truncated-CG inner step of the Newton
 * subproblem (solver.py:975-989) as one pass. */
int hfx_lincomb_dev_batch(int nops, size_t n1, int kind1, size_t n2, int kind2, const void* const* X1,
                          const void* const* Y1, void* const* out1, const void* const* X2, const void* const* Y2,
                          void* const* out2, const double* a, const int* a_idx, const double* b, const int* b_idx,
                          const double* dev_scalars, void* stream);
/* out_i = a' X_i + b' Y_i for two arrays in one launch (n2 = 0: the first
 * only; Y NULL: zero), a' = a * (*a_dev) when a_dev is given, else a (b' alike):
 * the updates of the Newton subproblem's truncated CG with coefficients that
 * stay on the device (solver.py:947-991 forms them as traced scalars).
 * out may alias X or Y. */
int hfx_lincomb_dev(size_t n1, int kind1, const void* X1, const void* Y1, void* out1, size_t n2, int kind2,
                    const void* X2, const void* Y2, void* out2, double a, const double* a_dev, double b,
                    const double* b_dev, void* stream);
/* Device scalars of the Steihaug truncated CG (solver.py:947-991), float64[16]. */
#define HFX_NT_RZ 0      /* r^T z of the current residual */
#define HFX_NT_DELTA 1   /* trust radius (set by the caller) */
#define HFX_NT_GRAD 2    /* gradient norm of the outer iterate (caller) */
#define HFX_NT_ALPHA 3
#define HFX_NT_TBD 4     /* step to the trust-region boundary along d */
#define HFX_NT_NEG 5     /* d^T H d <= 1e-14 d^T d */
#define HFX_NT_DONE 6    /* the loop has stopped: later steps change nothing */
#define HFX_NT_CV 7      /* v += CV d,  H v += CV H d */
#define HFX_NT_CR 8      /* r += CR H d */
#define HFX_NT_CDA 9     /* d = CDA z + CDB d */
#define HFX_NT_CDB 10
#define HFX_NT_NALPHA 11 /* -alpha */
#define HFX_NT_STEPS 12  /* inner steps taken */
#define HFX_NT_LEN 16

/* Trust-region Newton HVP pieces (solver.py:916-930):
 *   hfx_newton_dP:  dP = herm(Q (X (x) diff_p + diag(dp)) Q^dagger)
 *   hfx_newton_hvp: with R = F[dP] - h: HX = skew(diff_dp (x) Ft + diff_p (x) A) offdiag,
 *                   Hp = Re diag(A) + T/(pc(1-pc)) dp, A = [Ft, X] + Q^dagger R Q; part[k] = sum_b Hp
 *   hfx_newton_prec: ZX = RX / sqrt(gap^2 + st[LAM]^2), zp = rp pc(1-pc)/T; part[k] = sum_b zp
 *                    (solver.py:905-913, projection applied by the caller). */
int hfx_newton_dP(int dtype, int nb, size_t nk, const void* Q, const void* X, const double* p, const double* dp,
                  void* dP, void* stream);
int hfx_newton_hvp(int dtype, int nb, size_t nk, const void* Q, const void* R, const void* Ft, const void* X,
                   const double* p, const double* dp, double T, void* HX, double* Hp, double* part, void* stream);
int hfx_newton_prec(int dtype, int nb, size_t nk, const void* RX, const double* rp, const double* eps,
                    const double* p, const double* st, double T, void* ZX, double* zp, double* part, void* stream);

/* ---- reductions and scalar steps ------------------------------------------ */
#define HFX_OP_SUM 0
#define HFX_OP_GRAD 1
#define HFX_OP_BETA 2
#define HFX_OP_TAU0 3
#define HFX_OP_TRIAL 4
/* Steihaug truncated-CG scalar steps (solver.py:947-991) on nval / 2 reduced
 * vector pairs, matrix parts first, then the weighted occupation parts; st is
 * the loop's HFX_NT_* block.  NT2 starts the loop (rz = pair 0); NT0 follows the
 * Hessian-vector product (d^T H d, d^T d, v^T d, v^T v -> alpha, the boundary
 * step); NT1 the preconditioned residual (|v + alpha d|^2, |r - alpha H d|^2,
 * r^T z -> the stop test and the update coefficients, identity updates once
 * stopped). */
#define HFX_OP_NT0 5
#define HFX_OP_NT1 6
#define HFX_OP_NT2 7
#define HFX_OP_PARTIALS 8   /* hfx_reduce: the block partials only, summed by hfx_diis_coefficients_reduce */
/* Scratch (float64) for hfx_reduce / hfx_solve_mu. */
size_t hfx_solver_scratch_doubles(void);
/* out[v] = sum_k (w[k] if bit v of wmask) part[v][k] in a fixed order; then the
 * scalar step `op` on st.  params (10 doubles): {T, denom_scale, count,
 * max_step, shrink, tau_fallback, restart, last, omega0, restart_every}; the
 * HFX_OP_BETA step restarts (beta = 0) when `restart` is set or when
 * restart_every > 0 and st[HFX_ST_ITER] % restart_every == 0 (solver.py:517),
 * then advances st[HFX_ST_ITER]. */
int hfx_reduce(const double* part, int nval, size_t nk, const double* w, unsigned wmask, double* out,
               double* st, int op, const double* params, double* scratch, void* stream);

/* The weighted inner products of np <= 4 (matrix, vector) pairs of the
 * Newton subproblem (solver.py:947-991, the _jip of the (X, p) tangent
 * vectors): out[v] = sum_k (w_k if bit v of wmask) Re <X_v[k], Y_v[k]>_F and
 * out[np + v] = sum_k (w_k if bit np + v) <vx_v[k], vy_v[k]> (vectors of nb
 * reals per k), in a fixed order; then the scalar step op (HFX_OP_SUM or
 * HFX_OP_NT0..2 on st) as hfx_reduce.  X, Y: (nk, nb, nb) of dtype; w: (nk).
 * Two launches; scratch as hfx_reduce. */
int hfx_pair_dots(int dtype, int nb, size_t nk, int np, const void* const* X, const void* const* Y,
                  const double* const* vx, const double* const* vy, const double* w, unsigned wmask, double* out,
                  double* st, int op, double* scratch, void* stream);

#define HFX_MU_NEWTON_BRACKET 0
#define HFX_MU_BISECTION 1
#define HFX_MU_NEWTON 2
/* Chemical potential: Newton with bracket (solver.py:62-101) or safeguarded
 * Newton (utils.py:385-442) in one cooperative launch; bisection
 * (utils.py:353-382; maxiter levels, 54 in FP64 / 30 in FP32), twice with the
 * level shift of reference_scf.py:87-96 when level_shift != 0, as a
 * multisection of ordinary launches (three levels of the bisection tree per
 * pass, the same midpoints and tests).  Result to st[HFX_ST_MU] and/or
 * mu_out; occupations to p_out (sigmoid((mu-eps)/T), or fermidirac of the
 * shifted levels).  scratch: hfx_solver_scratch_doubles() doubles. */
int hfx_solve_mu(const double* eps, const double* w, size_t nk, int nb, double target, double T, double mu0,
                 int mu0_from_state, int maxiter, int mode, double level_shift, double* st, double* mu_out,
                 double* p_out, double* scratch, void* stream);

/* QR retraction (_qr_retract, solver.py:108-119): U(k) = Q of
 * I - tau[k] G(k) with the phases normalised so that R has a positive real
 * diagonal (the unique such Q); G and U (nk, nb, nb), U may alias G, tau (nk)
 * float64.  nb <= 32. */
int hfx_qr_retract(int dtype, int nb, size_t nk, const void* G, const double* tau, void* U, void* stream);

/* No counterpart in the reference.  hfx_solve_mu without the scratch argument,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the scratch
 * is a library-owned static device array of hfx_noalloc_scratch_doubles()
 * doubles (256 KiB) per GPU, held until the process ends; the call allocates
 * nothing.  nk * nb must be below 2^32 (HFX_EINVAL otherwise).  All calls on
 * one GPU share that array: calls that overlap on one GPU (several streams or
 * unordered host threads) race. */
size_t hfx_noalloc_scratch_doubles(void);
int hfx_solve_mu_noalloc(const double* eps, const double* w, size_t nk, int nb, double target, double T,
                         double mu0, int mu0_from_state, int maxiter, int mode, double level_shift, double* st,
                         double* mu_out, double* p_out, void* stream);

/* The safeguarded mu root of solver.py:62-101 as a fixed sequence of
 * maxiter + 2 ordinary launches that can be captured in a HIP graph: a
 * bracket pass, maxiter chained steps (each finishes the previous evaluation
 * -- a fourth-order step from N and its first three derivatives, inside the
 * bracket -- and evaluates at the new mu), and a finish that writes
 * st[HFX_ST_MU] / [MULO] / [MUHI] / [MUDONE] and the occupations
 * sigmoid((mu - eps) / T) to p_out.  Once converged, st[HFX_ST_MUDONE] is set
 * and the remaining steps only carry the state.  Same root as
 * hfx_solve_mu(HFX_MU_NEWTON_BRACKET) to its tolerance. */
int hfx_solve_mu_steps(const double* eps, const double* w, size_t nk, int nb, double target, double T,
                       double mu0, int mu0_from_state, int maxiter, double* st, double* p_out, double* scratch,
                       void* stream);

/* ---- batched solves (a density / field scan, no counterpart in the reference,
 * whose scan is a serial loop of solve calls: examples/multilayer_graphene_
 * density_scan.py:117-137).  nsolve solves of one mesh and band count are
 * stored back to back ((nsolve, nk1, nk2, nb, nb)), each with its own
 * HFX_ST_LEN state vector (st + HFX_ST_LEN * s); a solve whose
 * st[HFX_ST_FROZEN] is set is skipped by the reductions, mu, epilogue and the
 * _gj line-search kernels. ---- */

/* Exchange self-energy of nsolve stacked densities with one shared VR (one
 * launch; all nb^2 channels, or the upper triangle + mirror with herm). */
int hfx_selfenergy_batched(const hfx_plan* plan, int nsolve, const void* P, const void* VR, int vr_full, int herm,
                           void* Sigma, void* stream);
/* F = herm(h + Sigma) and E_s = sum_k w_k Re Tr[(h + Sigma/2) P] into
 * st[s][HFX_ST_E] (exchange-only Fock build of fock.py:76-97, 109-111);
 * w2d (nsolve, nk1, nk2); scratch of scratch_doubles >= 64 * nsolve doubles
 * (HFX_EINVAL when smaller). */
int hfx_fock_epilogue_batched(const hfx_plan* plan, int nsolve, const void* h, const void* Sigma, const void* P,
                              const double* w2d, void* F, double* st, double* scratch, size_t scratch_doubles,
                              void* stream);
/* hfx_fock_epilogue_batched reading Sigma from the slot planes of
 * hfx_selfenergy_planar (default channel set for herm); scratch of
 * scratch_doubles >= 256 * nsolve doubles (HFX_EINVAL when smaller). */
int hfx_fock_epilogue_planar_batched(const hfx_plan* plan, int nsolve, const void* h, const void* Sigt, int herm,
                                     const void* P, const double* w2d, void* F, double* st, double* scratch,
                                     size_t scratch_doubles, void* stream);
/* Pulay C-DIIS coefficients on the device (_diis_coefficients, reference_scf.py:99-121):
 * writes the new row brow (nh doubles, device) into row and column idx of the
 * device B matrix (nh x nh), solves the masked, max-normalised augmented
 * system over the first `count` slots (LU, partial pivoting) and writes
 * coef[0..7]: the solution when count >= 2 and the solve is regular and
 * finite, else the plain step e_idx.  nh <= 8. */
int hfx_diis_coefficients(const double* brow, double* B, int nh, int idx, int count, double* coef, void* stream);
/* hfx_reduce's sum and hfx_diis_coefficients in one launch: the block partials
 * that hfx_reduce(..., HFX_OP_PARTIALS, ..., scratch) left for nval values over
 * nk per-k sums are summed in hfx_reduce's own order into out[0..nval)
 * (bit-identical to the two-launch sum), and out[2 ..] is the new B row
 * (the DIIS polish's residual pass: |FP - PF|^2, |P - P_prev|^2, then the
 * row, reference_scf.py:110-121). */
int hfx_diis_coefficients_reduce(const double* scratch, size_t nk, int nval, double* out, double* B, int nh, int idx,
                                 int count, double* coef, void* stream);
/* The same for nsolve stacked solves: solve s reads brow + s * brow_stride and
 * owns B + s * nh * nh and coef + s * 8 (the batched polish of a scan). */
int hfx_diis_coefficients_batched(const double* brow, size_t brow_stride, double* B, int nsolve, int nh, int idx,
                                  int count, double* coef, void* stream);

/* Scratch (float64) for hfx_reduce_batched / hfx_solve_mu_batched over nseg
 * solves: block partials, then one completion-ticket word per solve at the
 * end.  The two entry points return HFX_EINVAL when scratch_doubles is
 * below this size; the scratch must not be shared with a buffer that other
 * launches write in the meantime (the tickets must stay zero between calls). */
size_t hfx_batch_scratch_doubles(int nseg);
/* hfx_reduce per solve (segments of seg_len k-points); with op == HFX_OP_SUM
 * and out_slot >= 0 the sums go to st[s][out_slot ...]. */
int hfx_reduce_batched(const double* part, int nval, size_t nk_total, int seg_len, const double* w, unsigned wmask,
                       double* st, int op, const double* params, int out_slot, double* scratch,
                       size_t scratch_doubles, void* stream);
/* hfx_solve_mu_steps per solve (targets[s] electrons, mu warm-started from
 * st[s][HFX_ST_MU]). */
int hfx_solve_mu_batched(const double* eps, const double* w, const double* targets, size_t nk_total, int seg_len,
                         int nb, double T, int maxiter, double* st, double* p_out, double* scratch,
                         size_t scratch_doubles, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HFX_H */
