/* librerdmft_gpu -- MI355X (gfx950) RI / Cholesky J/K Fock builds and RI AO->MO transforms.
 *
 * A C ABI over plain pointers and sizes. Every array is row-major, float64 or complex128 stored as
 * interleaved (re, im) doubles, and may live in host memory or in device memory of the handle's GPU
 * (the library detects which). Calls are synchronous: outputs are complete when a call returns.
 * Errors are reported as non-zero status codes; rr_last_error() returns the message, which for
 * shape errors is the message the reference throws as std::runtime_error.
 *
 * Reference call sites each entry point replaces are cited as file:line of marm314/ReRDMFT.
 */
#ifndef RERDMFT_GPU_H
#define RERDMFT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rr_context* rr_handle;

/* status codes */
#define RR_OK 0
#define RR_ERR_ARG 1      /* invalid argument */
#define RR_ERR_DIM 2      /* dimension mismatch (reference: std::runtime_error on shape checks) */
#define RR_ERR_HIP 3      /* HIP runtime error */
#define RR_ERR_SOLVER 4   /* eigensolver failure (reference: LAPACK non-convergence) */
#define RR_ERR_STATE 5    /* missing tensor slot / precision copy */
#define RR_ERR_INTERNAL 6

/* arithmetic of a contraction: operands stored in this type, matrix cores accumulate in
 * fp64 (RR_FP64) or fp32 (RR_FP32, RR_BF16) */
#define RR_FP64 0
#define RR_FP32 1
#define RR_BF16 2

/* copies of a 3-index tensor kept resident in HBM */
#define RR_COPY_FP64 1u
#define RR_COPY_FP32 2u
#define RR_COPY_BF16 4u

/* ---- context ---------------------------------------------------------------------------- */
int rr_create(int device, rr_handle* out);
int rr_destroy(rr_handle h);
const char* rr_last_error(rr_handle h);
/* Select the GEMM tile configuration: -1 per-shape automatic choice; 0..7 a register-direct
 * configuration for every launch; 100..142 an LDS-staged configuration for the hot real GEMMs
 * (100..106 for every operand type, 107..125 and 130..136 for bf16 and 126..129, 137..138 for
 * fp64 / fp32 operands only; 139..141 the row-merged persistent bf16 half-transform and 142 the
 * persistent split-K bf16 rank-k update, each used where the product's shape fits it; other
 * operand types and shapes keep their default).
 * Results do not depend on the choice (every element is summed in the same order). A change drops
 * the captured graph of rr_ri_jk_multi, which holds the kernels chosen when it was captured. */
int rr_set_tile_config(rr_handle h, int cfg);
/* How the density-taking builds (rr_nonrel_fock_ri, rr_x2c_fock_ri, rr_rkb_fock_ri) find the
 * occupied set, in place of the reference's eigendecomposition of X^H (S P S) X (e.g.
 * X2C_DHF/X2C_FockMatrixRi.cpp:65-95); no counterpart in the reference. method 0 (default): spectral
 * projector onto the eigenvalues above 1/2 by a Newton-Schulz sign iteration and a factor of the
 * occupied part, falling back to the eigensolver when the iteration does not settle (an eigenvalue at
 * or near 1/2); 1: eigensolver only; -1: leave unchanged. Both give the same Fock matrix (the
 * exchange depends only on sum over occupied of occupation v v^H). The counters receive the number of
 * recoveries each method has finished on this handle (either pointer may be NULL). */
int rr_set_recovery_method(rr_handle h, int method, long long* n_projector, long long* n_eigensolver);
/* Per-phase device time (HIP events on the library stream) accumulated since the last reset,
 * indexed 8 * precision + phase (n up to 24): phase 0 Coulomb gamma contraction, 1 Coulomb J
 * stream, 2 exchange half-transform GEMM, 3 exchange rank-k GEMM, 4 assembly, 5 recovery;
 * counts[i] = launches timed. enable=0 turns the event timing off (default off). */
int rr_timing(rr_handle h, int enable, int reset, double* ms, long long* counts, int n);

/* ---- 3-index tensors (AO_ints/ThreeCenterIntegrals.h:42-70 layout: (n_aux, n*n), row P is the
 * n x n block (mu nu|P)); also the Cholesky vectors of Utils/AoCholesky.h:23-41 ------------- */
int rr_tensor_upload(rr_handle h, int slot, const double* eri3, size_t n_aux, size_t n, unsigned copies);
/* Synthetic symmetric tensor generated on the device: sigma * N(0,1) * exp(-|mu-nu|/decay). */
int rr_tensor_generate(rr_handle h, int slot, size_t n_aux, size_t n, double sigma, double decay, uint64_t seed,
                       unsigned copies);
int rr_tensor_info(rr_handle h, int slot, size_t* n_aux, size_t* n, unsigned* copies);
int rr_tensor_release(rr_handle h, int slot);

/* ---- NON_REL ------------------------------------------------------------------------------ */
/* nonRelFockMatrix(h, RiNonRelEriSource, P): src/NON_REL/NonRelFockMatrixRi.cpp:27-100
 * (declared NonRelFockMatrixRi.h:39-40). F = h + J - K/2 with the reference's occupied-orbital
 * recovery from P (eigenvalues of X^T (S P S) X / 2 above 0.5). j_out / k_out may be NULL.
 * On a slot holding Cholesky vectors (n_chol in place of n_aux) this is the factor form of
 * nonRelFockMatrix(h, AoCholesky, P) (src/Utils/AoCholesky.cpp:108-149) for an SCF density:
 * K = sum_L (B_L C)(B_L C)^T equals sum_L B_L P B_L when P = C occ C^T with C^T S C = I. */
int rr_nonrel_fock_ri(rr_handle h, int slot, int precision, size_t nl, const double* h_core, const double* x_large,
                      size_t nx, const double* overlap, const double* density, double* fock, double* j_out,
                      double* k_out);

/* Factor-form real J/K without recovery (the exchange loop of NonRelFockMatrixRi.cpp:74-94 with a
 * caller-supplied factor): J = sum_P tr(B_P D) B_P, K = sum_P B_P^T (L R^T) B_P with L, R n x r
 * (rfac NULL: R = L). Either of j_out / k_out may be NULL. density NULL with j_out set means
 * D = L R^T: the Coulomb contraction is then taken from the half-transformed factors inside the
 * exchange GEMM, saving one pass over the tensor. */
int rr_ri_jk_factor(rr_handle h, int slot, int precision, size_t n, const double* density, const double* lfac,
                    const double* rfac, size_t r, double* j_out, double* k_out);

/* Allocation-free variant of rr_ri_jk_factor for D = L L^T (density NULL, rfac NULL; J from the
 * half-transformed factor). No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid it). It
 * calls no hipMalloc / hipFree and no allocating library routine: every workspace lives in one
 * static 160 MiB scratch region per device (held until the process ends). lfac, j_out and k_out
 * must be device memory; a shape whose workspace exceeds the region is rejected with
 * RR_ERR_ARG before any launch. Results equal rr_ri_jk_factor's. Every call on a device shares that
 * device's region, so calls that overlap on one GPU (other threads, other handles) race. */
int rr_ri_jk_factor_noalloc(rr_handle h, int slot, int precision, size_t n, const double* lfac, size_t r,
                            double* j_out, double* k_out);

/* Many densities sharing the tensor and the left factor (no reference counterpart; the trial
 * densities of the orbital-Hessian sigma products, D_k = L R_k^T with L the occupied orbitals):
 * for k < n_dens, J_k = sum_P tr(B_P D_k) B_P and K_k = sum_P B_P^T (L R_k^T) B_P, as n_dens calls of
 * rr_ri_jk_factor(density = NULL, lfac, rfac + k n r) would give. lfac is n x r, rfac n_dens
 * consecutive n x r blocks, j_out / k_out n_dens consecutive n x n blocks (either may be NULL).
 * B L is formed once; the first call for a shape runs eagerly and captures a HIP graph of the whole
 * build, which later calls of the same shape replay (replayed, optional, reports it). */
int rr_ri_jk_multi(rr_handle h, int slot, int precision, size_t n, size_t r, const double* lfac, size_t n_dens,
                   const double* rfac, double* j_out, double* k_out, int* replayed);

/* Incremental build for D = 2 C C^T (no reference counterpart; replaces a full call of
 * NonRelFockMatrixRi.cpp:27-100 between two SCF iterations): dJ = J(D_new - D_old),
 * dK = K(D_new - D_old) from the aligned-difference factorization of D_new - D_old, so the
 * low-precision error is proportional to |D_new - D_old|. */
int rr_ri_jk_delta(rr_handle h, int slot, int precision, size_t n, size_t n_occ, const double* c_new,
                   const double* c_old, double* dj_out, double* dk_out);

/* ---- X2C ---------------------------------------------------------------------------------- */
/* x2cFockMatrix(h, RiX2cEriSource, P): src/X2C_DHF/X2C_FockMatrixRi.cpp:35-162 (declared
 * X2C_FockMatrixRi.h:38-40). Complex 2nl x 2nl arrays; x_large is 2nl x nx. On a slot holding
 * Cholesky vectors this is the factor form of x2cFockMatrix(h, AoCholesky, P)
 * (src/Utils/AoCholesky.cpp:151-217) for an SCF density (P = C occ C^H with C^H S C = I). */
int rr_x2c_fock_ri(rr_handle h, int slot, int precision, size_t nl, const double* h_x2c, const double* x_large,
                   size_t nx, const double* overlap, const double* density, double* fock);

/* ---- 4c RKB (UKB blocks) ------------------------------------------------------------------- */
/* riFockTwoElectronDirect: src/UKB/UkbFockMatrixRi.cpp:35-183 (declared UkbFockMatrixRi.h:70-75).
 * slot_ll / slot_ss hold eri3_LL (n_aux, nl^2) and eri3_SS (n_aux, ns^2) sharing one aux index. */
int rr_ri_fock_two_electron_direct(rr_handle h, int slot_ll, int slot_ss, int precision, size_t nl, size_t ns,
                                   const double* density_ukb, const double* c_occ_ukb, size_t n_occ,
                                   const double* occupations, double* g_out);
/* rkbFockMatrix(h, RiDirectEriSource, P): src/UKB/UkbFockMatrixRi.cpp:185-216 (declared
 * UkbFockMatrixRi.h:96-98). v_total is n_ukb x n_rkb, s_full n_rkb x n_rkb, x_full n_rkb x nx. */
int rr_rkb_fock_ri(rr_handle h, int slot_ll, int slot_ss, int precision, size_t nl, size_t ns, const double* h_rkb,
                   const double* v_total, size_t n_rkb, const double* s_full, const double* x_full, size_t nx,
                   const double* density, double* fock);

/* ---- MO-basis RI tensors ------------------------------------------------------------------- */
/* buildRiMoEri: src/UKB/RiMoEri.cpp:64-86 (+ transformBlock :20-62). c_ukb is n_ukb x n_total
 * complex; out is (n_aux, n_total^2) complex, B(P,p,q) = sum conj(C(mu,p)) C(nu,q) (mu nu|P).
 * ns = 0 gives the X2C 2-block transform with slot_ss unused. out_precision RR_FP64 writes
 * complex128, RR_FP32 writes complex64 (interleaved floats). */
int rr_build_ri_mo_eri(rr_handle h, int slot_ll, int slot_ss, int precision, int out_precision, size_t nl, size_t ns,
                       const double* c_ukb, size_t n_total, void* out);
/* RiMoEri::rotated: src/UKB/RiMoEri.cpp:88-110. B'_P = U^H B_P U for b (n_aux, n^2) complex. */
int rr_ri_mo_eri_rotated(rr_handle h, int precision, size_t n_total, size_t n_aux, const double* b, const double* u,
                         double* out);
/* RiMoEri::positiveBlock: src/UKB/RiMoEri.cpp:112-121. */
int rr_ri_mo_eri_positive_block(rr_handle h, size_t n_total, size_t n_aux, const double* b, size_t off, double* out);
/* buildRiNonRelSpinMoEri: src/NON_REL/NonRelSpinRiMoEri.cpp:13-43. B^MO_P = C^T B_P C (real);
 * c_spatial is n x n_mo (n_mo = n in the reference). out (n_aux, n_mo^2). */
int rr_build_ri_nonrel_spin_mo_eri(rr_handle h, int slot, int precision, size_t n, const double* c_spatial,
                                   size_t n_mo, double* out);
/* RiNonRelSpinMoEri::rotated: src/NON_REL/NonRelSpinRiMoEri.cpp:45-95 (u is 2n x 2n). */
int rr_ri_nonrel_spin_mo_eri_rotated(rr_handle h, int precision, size_t n_spatial, size_t n_aux, const double* b,
                                     const double* u, double* out);
/* RiNonRelSpinMoEri::embeddedB: src/NON_REL/NonRelSpinRiMoEri.h:77-89. */
int rr_ri_nonrel_spin_embedded_b(rr_handle h, size_t n_spatial, size_t n_aux, const double* b, double* out);

/* ---- Cholesky vectors (slot uploaded from AoCholesky::vectors, (n_chol, n*n)) ---------------- */
/* nonRelFockMatrix(h, AoCholesky, P): src/Utils/AoCholesky.cpp:108-149 (K = sum_L B_L P B_L, any
 * symmetric P). For an SCF density rr_nonrel_fock_ri on the same slot gives the factor form. */
int rr_nonrel_fock_cholesky(rr_handle h, int slot, int precision, size_t n, const double* h_core, const double* p,
                            double* fock);
/* x2cFockMatrix(h, AoCholesky, P): src/Utils/AoCholesky.cpp:151-217 (density form, any Hermitian
 * P). For an SCF density rr_x2c_fock_ri on the same slot gives the factor form. */
int rr_x2c_fock_cholesky(rr_handle h, int slot, int precision, size_t n, const double* h_x2c, const double* p,
                         double* fock);
/* aoCholeskyToMoSpinOrbital (spatial block C^T B_L C): src/Utils/AoCholesky.cpp:219-243. */
int rr_cholesky_to_mo_spatial(rr_handle h, int slot, int precision, size_t n, const double* c, size_t n_mo,
                              double* out);
/* aoCholeskyToMoSpinor: src/Utils/AoCholesky.cpp:270-295 (out W_L = (sum_s C_s^H B_L C_s)^T). */
int rr_cholesky_to_mo_spinor(rr_handle h, int slot, int precision, size_t n, const double* c, size_t n_mo,
                             double* out);

/* ---- Functional stage: RI generalized Fock and sigma-vector engine on MO RI tensors ---------
 * b is an MO RI tensor (n_aux, n*n), B(P,p,q) at [P*n*n + p*n + q], real or complex interleaved
 * (is_complex), with eri(a,b,c,d) = sum_P B(P,a,c) B(P,b,d). Couplings two_rdm_* are real n x n;
 * pair_of (optional, length n) is the pair-partner involution. Host or device pointers. */

/* riHxFockGemm: src/Hessian_opt/RiGemmFock.cpp:53-115, behind hartreeExchangeFockMatrix /
 * jkOnlyFockMatrix for RiMoEri and RiNonRelSpinMoEri (:119-163; pass embeddedB for the latter).
 * F(p,q) = occ_q h(q,p) + sum_s [h2(s,q) eri(q,s,p,s) - x2(q,s) eri(q,s,s,p)]
 *          + [pair_of] 2 sum_s eri(s,sbar,qbar,p) (l1(s,qbar) + l2(s,q)). */
int rr_ri_hx_fock_gemm(rr_handle h, int is_complex, size_t n, size_t n_aux, const double* b, const double* h_mo,
                       const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                       const size_t* pair_of, const double* two_rdm_l1, const double* two_rdm_l2, double* fock);

/* buildRiSigmaHessianToolkit: src/Hessian_opt/RiSigmaHessianToolkit.cpp:25-56 (struct
 * RiSigmaHessianToolkit.h:44-61). kappa real n x n. Outputs (any may be NULL): kb1, kb2 (n_aux, n*n);
 * bdiag, kb1diag, kb2diag (n_aux, n). */
int rr_ri_sigma_toolkit(rr_handle h, int is_complex, size_t n, size_t n_aux, const double* b, const double* kappa,
                        double* kb1, double* kb2, double* bdiag, double* kb1diag, double* kb2diag);

/* Operand of a pair contraction (risigma::PairOp, src/Hessian_opt/RiSigmaGemm.h:23-45):
 *   value(P,t,x) = [s ? s(P, idx[t*n+x]) : 1] * [w ? w[t*n+x] : 1] * [row ? row(P,t) : 1],
 * right-multiplied by post (real n x n) when given. s (n_aux, n*n) and row (n_aux, n) have the
 * element type of the call; idx (uint32 n*n) is required with s. */
typedef struct rr_pair_op {
  const double* s;
  const uint32_t* idx;
  const double* w;
  const double* row;
  const double* post;
} rr_pair_op;

/* pairContract: src/Hessian_opt/RiSigmaGemm.cpp:76-101.
 * target(x,y) += sign * sum_{P,t} F(P,t,x) G(P,t,y); target n x n (in/out). */
int rr_ri_pair_contract(rr_handle h, int is_complex, size_t n, size_t n_aux, double sign, const rr_pair_op* f,
                        const rr_pair_op* g, double* target);
/* diagContract: src/Hessian_opt/RiSigmaGemm.cpp:103-121.
 * target(a,b) += sum_P S(P, b*n+a) (Ua(P,a) + Ub(P,b)); ua/ub (n_aux, n) may be NULL. */
int rr_ri_diag_contract(rr_handle h, int is_complex, size_t n, size_t n_aux, const double* s, const double* ua,
                        const double* ub, double* target);
/* accumU: src/Hessian_opt/RiSigmaGemm.cpp:123-132. U(P,x) += alpha sum_t D[P*n+t] Wm(x,t);
 * d, u (n_aux, n) of the call's element type, wm real n x n. */
int rr_ri_accum_u(rr_handle h, int is_complex, size_t n, size_t n_aux, double alpha, const double* d, const double* wm,
                  double* u);

/* Sigma-vector orbital-Hessian products w = H v. pairs: n_pairs (p,q) index pairs, flattened.
 * jkOnlySigmaHessianVectorRi: src/Hessian_opt/JkOnlySigmaHessianVectorRi.cpp:504-523 (real). */
int rr_jk_only_sigma_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b, size_t n_aux,
                                       const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                                       size_t n_pairs, const size_t* pairs, const double* v, double* w);
/* jkOnlySigmaJointHessianVectorRi: JkOnlySigmaHessianVectorRi.cpp:525-573 (complex h, b;
 * v and w have 2*n_pairs entries [t; y]). */
int rr_jk_only_sigma_joint_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b, size_t n_aux,
                                             const double* occupations, const double* two_rdm_h,
                                             const double* two_rdm_x, size_t n_pairs, const size_t* pairs,
                                             const double* v, double* w);
/* hartreeExchangeSigmaHessianVectorRi: src/Hessian_opt/HartreeExchangeSigmaHessianVectorRi.cpp:442-484. */
int rr_hartree_exchange_sigma_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b,
                                                size_t n_aux, const double* occupations, const double* two_rdm_h,
                                                const double* two_rdm_x, const double* fock, const size_t* pair_of,
                                                const double* two_rdm_l1, const double* two_rdm_l2, size_t n_pairs,
                                                const size_t* pairs, const double* v, double* w);
/* hartreeExchangeSigmaJointHessianVectorRi: HartreeExchangeSigmaHessianVectorRi.cpp:486-539. */
int rr_hartree_exchange_sigma_joint_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b,
                                                      size_t n_aux, const double* occupations, const double* two_rdm_h,
                                                      const double* two_rdm_x, const double* fock,
                                                      const size_t* pair_of, const double* two_rdm_l1,
                                                      const double* two_rdm_l2, size_t n_pairs, const size_t* pairs,
                                                      const double* v, double* w);

/* ---- One-time RI tensor preparation (FP64) -----------------------------------------------------
 * riThreeCenterTensor's contraction: src/AO_ints/ThreeCenterIntegrals.cpp:219-236.
 * out (n_aux_kept, n*n) = metric_x^T (n_aux_raw x n_aux_kept)^T . raw (n_aux_raw, n*n); the raw
 * integrals themselves come from the caller's integral engine. */
int rr_ri_three_center_tensor(rr_handle h, size_t n, size_t n_aux_raw, const double* raw, const double* metric_x,
                              size_t n_aux_kept, double* out);
/* transformRiThreeCenterAoLegs: src/AO_ints/ThreeCenterIntegrals.cpp:191-217.
 * eri3_out[P] = t^T eri3_in[P] t, t is n_ao x n_out. */
int rr_transform_ri_three_center_ao_legs(rr_handle h, size_t n_aux, size_t n_ao, const double* eri3_in,
                                         const double* t, size_t n_out, double* eri3_out);
/* canonicalOrthogonalize (the auxiliary-metric step of auxMetricOrthogonalization,
 * src/AO_ints/ThreeCenterIntegrals.cpp:186-189): src/Linear_Algebra/LinearAlgebra.cpp:267-310.
 * x receives n x n_kept (row-major; the caller provides n*n doubles); n_kept = eigenvalues > threshold.
 * With nothing dropped x is the symmetric S^-1/2. */
int rr_canonical_orthogonalize(rr_handle h, size_t n, const double* s, double threshold, double* x, size_t* n_kept);

/* Fused functional-stage Fock builds (one call per orbital-optimisation step; the MO tensor stays
 * on the device). The functional enters through the couplings two_rdm_h/x (and pair terms), which
 * the caller computes (jkHartreeCoupling, PNOF/pCCD 2-RDM builders).
 * riJkOnlyFockMatrix / riPnofFockMatrix body: src/Hessian_opt/RiJkOnlyFock.cpp:14-33,
 * RiPnofFock.cpp:14-30 (C_ukb = v_total c_rkb, buildRiMoEri, h_mo = C^H h C, riHxFockGemm).
 * h_rkb n_rkb x n_rkb, v_total (2nl+2ns) x n_rkb, c_rkb n_rkb x n_mo (complex); fock n_mo x n_mo. */
int rr_ri_mo_fock_rkb(rr_handle h, int slot_ll, int slot_ss, size_t nl, size_t ns, const double* h_rkb,
                      const double* v_total, size_t n_rkb, const double* c_rkb, size_t n_mo, const double* occupations,
                      const double* two_rdm_h, const double* two_rdm_x, const size_t* pair_of, const double* two_rdm_l1,
                      const double* two_rdm_l2, double* fock);
/* riJkOnlyFockMatrixNonRel / riPnofFockMatrixNonRel body: src/Hessian_opt/RiNonRelFullOptFock.cpp:53-98.
 * h_core_ao n x n, c_current n x 2n (real); couplings and fock 2n x 2n. */
int rr_ri_mo_fock_nonrel(rr_handle h, int slot, size_t n_spatial, const double* h_core_ao, const double* c_current,
                         const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                         const size_t* pair_of, const double* two_rdm_l1, const double* two_rdm_l2, double* fock);

/* ---- 4c RKB Cholesky vectors (RkbCholesky, src/C4_DHF/RkbCholesky.h) ---------------------------
 * A set holds n_chol vectors, each with components large (real nl x nl), small[0], small[1],
 * small_cross[0], small_cross[1] (complex nl x nl); arrays are (n_chol, nl*nl) row-major. */
int rr_rkb_cholesky_upload(rr_handle h, int set, size_t n_large, size_t n_chol, const double* large,
                           const double* small0, const double* small1, const double* cross_fwd,
                           const double* cross_bwd);
int rr_rkb_cholesky_release(rr_handle h, int set);
/* rkbFockMatrix(h_rkb, RkbCholesky, P): src/C4_DHF/RkbCholesky.cpp:352-448 (n = 4*nl). */
int rr_rkb_fock_cholesky(rr_handle h, int set, size_t n, const double* h_rkb, const double* p, double* fock);
/* rkbFockMatrix(h_rkb, RkbCholesky, P) in factor form: the exchange of src/C4_DHF/RkbCholesky.cpp:381-432
 * (K_fg = sum_L B^f_L P_fg B^g_L) as K = sum_L W_L W_L^H with W^f_L = B^f_L F_f, F the occupied factor
 * recovered from P in the RKB metric as the RI build does (UkbFockMatrixRi.cpp:185-211: eigenvectors
 * of X^H (S P S) X above 1/2, weighted by sqrt(eigenvalue); s_full n x n, x_full n x nx complex).
 * Equal to rr_rkb_fock_cholesky when P = F F^H (an SCF density); the Coulomb part, with its
 * cross-flavor terms, is taken from P as there. Needs Hermitian large/small vectors (checked at
 * upload; RR_ERR_ARG otherwise). */
int rr_rkb_fock_cholesky_factor(rr_handle h, int set, size_t n, const double* h_rkb, const double* s_full,
                                const double* x_full, size_t nx, const double* p, double* fock);
/* rkbCholeskyToMo: src/C4_DHF/RkbCholesky.cpp:450-523. c_mo is (4 nl) x n_mo complex. With
 * n_negative = 0, out receives n_chol vectors of (n_mo x n_mo); otherwise the recompressed vectors
 * (RkbCholesky.cpp:244-283, eigenvalue cut threshold / n_chol) embedded in n_mo x n_mo. capacity =
 * the number of vectors out can hold; n_out = the number written. */
int rr_rkb_cholesky_to_mo(rr_handle h, int set, const double* c_mo, size_t n_mo, size_t n_negative, double threshold,
                          double* out, size_t capacity, size_t* n_out);

/* ---- Batches of independent molecules (no reference counterpart: many small systems per launch) --
 * A batch keeps n_mol RI tensors of different sizes resident; rr_batch_jk_factor builds J and K of
 * D_m = L_m L_m^T for every molecule (L_m n_m x r_m) in one HIP-graph launch (captured on the first
 * call for a given set of ranks and precision, replayed afterwards). replayed (optional) reports
 * whether the call replayed a captured graph. Per-molecule results equal rr_ri_jk_factor's. */
typedef struct rr_batch* rr_batch_t;
int rr_batch_create(rr_handle h, size_t n_mol, const size_t* n_aux, const size_t* n, const double* const* eri3,
                    unsigned copies, rr_batch_t* out);
int rr_batch_destroy(rr_handle h, rr_batch_t batch);
int rr_batch_jk_factor(rr_handle h, rr_batch_t batch, int precision, const size_t* r, const double* const* lfac,
                       double* const* j_out, double* const* k_out, int* replayed);

/* ---- SCF starting-density network (no reference counterpart; its fallback is the core-Hamiltonian
 * guess of main.cpp:5248-5250) -------------------------------------------------------------------
 * A fully connected FP32 network n_in -> n_layers x (n_hidden, tanh) -> n_out (linear), trained on
 * the device. Parameters are laid out per layer as W [fan_out][fan_in] row-major, then b [fan_out].
 * Inputs and targets are row-major double arrays [batch][n_in] / [batch][n_out] (host or device).
 * The loss is 1/(2 batch n_out) sum (y - target)^2; rr_mlp_train_step applies one Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8) and returns the loss before it. */
typedef struct rr_mlp* rr_mlp_t;
int rr_mlp_create(rr_handle h, int n_in, int n_hidden, int n_layers, int n_out, unsigned long long seed, rr_mlp_t* out);
int rr_mlp_destroy(rr_handle h, rr_mlp_t mlp);
int rr_mlp_num_params(rr_handle h, rr_mlp_t mlp, size_t* n);
int rr_mlp_get_params(rr_handle h, rr_mlp_t mlp, float* out);
int rr_mlp_set_params(rr_handle h, rr_mlp_t mlp, const float* in);
int rr_mlp_forward(rr_handle h, rr_mlp_t mlp, size_t batch, const double* x, double* y);
int rr_mlp_loss_grad(rr_handle h, rr_mlp_t mlp, size_t batch, const double* x, const double* target, double* loss,
                     float* grads);
int rr_mlp_train_step(rr_handle h, rr_mlp_t mlp, size_t batch, const double* x, const double* target, double lr,
                      double* loss);

/* ---- SCF DIIS on the device: scfCommutatorError and Diis<T>::extrapolate
 * (Utils/DIIS.h:13-70, Utils/DIIS.cpp:51-138) ----------------------------------------------------
 * rr_scf_commutator_error: e = F (P S) - (S P) F for row-major n x n real (is_complex = 0) or
 * interleaved complex matrices. rr_diis_extrapolate stores (error, value) -- n_error / n_value doubles
 * each, a complex entry counting twice -- keeping the newest max_vectors pairs in device memory, and
 * writes sum_i w_i value_i with the weights minimizing |sum_i w_i e_i|^2 under sum_i w_i = 1 (B_ij =
 * Re <e_i|e_j>, scaled by its largest diagonal; pivot ratio below 1e-14 drops the oldest pair);
 * with fewer than two usable pairs, max_vectors < 2 or a zero error the value is copied unchanged and
 * *n_weights = 0. weights (up to max_vectors doubles, newest first) and n_weights may be NULL.
 * max_vectors is at most 16. Errors: "Diis::extrapolate: shapes differ from the stored history". */
typedef struct rr_diis* rr_diis_t;
int rr_scf_commutator_error(rr_handle h, int is_complex, size_t n, const double* fock, const double* density,
                            const double* overlap, double* error);
int rr_diis_create(rr_handle h, size_t max_vectors, rr_diis_t* out);
int rr_diis_destroy(rr_handle h, rr_diis_t diis);
int rr_diis_clear(rr_handle h, rr_diis_t diis);
int rr_diis_extrapolate(rr_handle h, rr_diis_t diis, size_t n_error, const double* error, size_t n_value,
                        const double* value, double* out, double* weights, size_t* n_weights);

/* ---- CASSCF RI contractions (MCSCF/CasFock.cpp:70-150, RI backends :244-262) ------------------
 * b: the RI MO tensor (n_aux, n*n), row P = the n x n matrix B_P (Hermitian), real or interleaved
 * complex (is_complex); <ab|cd> = sum_P B(P, a n + c) B(P, b n + d). Host or device arrays.
 * rr_cas_two_electron_operator_ri: g (n x n) = G[rho]_rp = sum_st rho_st (<sr|tp> - <rs|tp>) for a
 *   Hermitian rho (riTwoElectronOperator, CasFock.cpp:73-113; rho's eigenvalues |w| <= 1e-14 dropped).
 * rr_cas_active_integrals_ri: out (a^2 x a^2) = (tu|vw) = sum_P B(P, act_t n + act_u) B(P, act_v n + act_w)
 *   (riActiveIntegrals, :136-148).
 * rr_cas_active_cumulant_fock_ri: f (n x a), F(p,q) = sum_r sum_P B(P, act_r n + p)
 *   sum_(s,t) B(P, act_s n + act_t) lambda[((s a + t) a + r) a + q] (riActiveCumulantFock, :115-134). */
int rr_cas_two_electron_operator_ri(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b,
                                    const double* rho, double* g);
int rr_cas_active_integrals_ri(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b, size_t n_act,
                               const size_t* act, double* out);
int rr_cas_active_cumulant_fock_ri(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b, size_t n_act,
                                   const size_t* act, const double* lambda, double* f);

/* UCCSD integrals of an RI MO tensor (CC/Uccsd.cpp:128-180, makeCcIntegrals(RiMoEri) :181-184):
 * out[p][q][r][s] = <P_p Q_q | R_r S_s> = sum_aux B(aux, P_p n + R_r) B(aux, Q_q n + S_s) for the index
 * lists p (np), q (nq), r (nr), s (ns) (physicsBlock; slab(p, list) is the block ({p}, list, list,
 * list)). The NON_REL spin-orbital source (makeCcIntegrals(RiNonRelSpinMoEri)) is the same call on
 * the spin-embedded tensor of rr_ri_nonrel_spin_embedded_b. */
int rr_cc_physics_block(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b, size_t np, const size_t* p,
                        size_t nq, const size_t* q, size_t nr, const size_t* r, size_t ns, const size_t* s, double* out);

#ifdef __cplusplus
}
#endif

#endif /* RERDMFT_GPU_H */
