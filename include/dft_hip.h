/*
 * dft_hip.h - C ABI of libdft_hip.so, the MI355X (gfx950) Gaussian-integral
 * and direct Fock-build library.
 *
 * Conventions
 *  - Every function returns int: 0 on success, negative on error
 *    (DFT_EINVAL bad argument, DFT_EUNSUPPORTED angular momentum or mode not
 *    built, DFT_EHIP a HIP runtime error). The message of the last error on
 *    the calling thread is returned by dft_last_error(). No C++ exception
 *    crosses this boundary.
 *  - Plan-creation inputs are host arrays. Per-iteration inputs and outputs
 *    (densities, J, K, S, T, V, gradients) are caller-owned device pointers,
 *    row-major nbf x nbf FP64 in the caller's Cartesian AO order.
 *  - stream is a hipStream_t (NULL = the null stream). Per-iteration calls
 *    only enqueue work: they neither allocate nor synchronise, so a sequence
 *    of them can be captured into a hipGraph.
 *  - J/K are deterministic: no atomics, a fixed-order reduction.
 *  - Distinct plans may be used concurrently from different threads/streams;
 *    one plan must not be used by two streams at once (it owns its staging).
 */
#ifndef DFT_HIP_H
#define DFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFT_OK 0
#define DFT_EINVAL -1
#define DFT_EUNSUPPORTED -2
#define DFT_EHIP -3

/* plan flags */
#define DFT_PLAN_UHF 1     /* K for two densities (alpha, beta) */
#define DFT_PLAN_NOSPLIT 2 /* one lane per quartet in every class (no primitive split
                              of classes with few quartets; makes the summation
                              order of a molecule the same in single and batch plans) */
#define DFT_PLAN_SERIAL 4  /* all class launches of a Fock build on the caller's stream
                              (no concurrent launches; for per-kernel profiling) */
#define DFT_PLAN_MFMA 16    /* heavy d class combinations (FP64 part) on the matrix cores
                              (MFMA products for the ket and bra contractions) instead of
                              the VALU column kernel */
#define DFT_PLAN_STATIC_WORKSPACE 8 /* allocate no Fock-build workspace: the plan's J/K are
                              built only by dft_jk_static (dft_jk, graphs and the tuner
                              reject it) */

/* Size of the library's static device scratch used by dft_jk_static. */
#define DFT_STATIC_SCRATCH_BYTES (32LL << 20)

/* precision modes of dft_jk */
#define DFT_FP64 0  /* every quartet in FP64 */
#define DFT_MIXED 1 /* quartets with Schwarz bound Q_ab Q_cd < fp32_tau in FP32 */

typedef struct dft_plan_s* dft_plan;

/* Version of the library, as major*10000 + minor*100 + patch. */
int dft_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* dft_last_error(void);

/*
 * Build the integral plan of one molecule: primitive-pair records (pair
 * pre-pass; FP64 and two FP32 record sets, each primitive pair with its own
 * Schwarz factor), Schwarz bounds Q_ab = sqrt(max |(ab|ab)|), the screened,
 * angular-momentum-class-sorted quartet lists (FP64 and FP32 parts) and the
 * fixed-order reduction map.
 *
 * Replaces the one-time set-up of the reference: shell construction and
 * normalisation (shelled.py:934-952), nuclear_kernel pair data
 * (shelled.py:350-450 launched from :1000-1058) and the quartet enumeration
 * of calc_eri (shelled.py:634-640).
 *
 *  nsh                 number of segmented shells (one coefficient row each)
 *  shell_l[nsh]        angular momentum (0 = s, 1 = p, 2 = d, 3 = f)
 *  shell_nprim[nsh]    primitives per shell
 *  shell_ao[nsh]       first Cartesian AO of the shell (components follow in
 *                      the order x, y, z for p; xx, xy, yy, xz, yz, zz for d;
 *                      xxx, xxy, xyy, yyy, xxz, xyz, yyz, xzz, yzz, zzz for
 *                      f, as combinations() of shelled.py:15-23)
 *  shell_xyz[3*nsh]    shell centre (bohr)
 *  prim_exp[], prim_coef[]  concatenated per shell: exponents and raw
 *                      contraction coefficients; primitive normalisation
 *                      (shelled.py:95-101) is applied inside
 *  natom, atom_xyz[3*natom], atom_charge[natom]  nuclei for V
 *  nbf                 number of Cartesian AOs
 *  schwarz_tau         quartets with Q_ab Q_cd < schwarz_tau are skipped
 *  fp32_tau            quartets with Q_ab Q_cd < fp32_tau form the FP32 part
 *                      (class combinations with an f shell stay FP64)
 *  flags               DFT_PLAN_UHF | DFT_PLAN_NOSPLIT | DFT_PLAN_SERIAL |
 *                      DFT_PLAN_STATIC_WORKSPACE | DFT_PLAN_MFMA or 0
 */
int dft_plan_create(int nsh, const int32_t* shell_l, const int32_t* shell_nprim, const int32_t* shell_ao,
                    const double* shell_xyz, const double* prim_exp, const double* prim_coef, int natom,
                    const double* atom_xyz, const double* atom_charge, int nbf, double schwarz_tau,
                    double fp32_tau, int flags, dft_plan* out);

/*
 * Plan of a batch of independent molecules: molecule m owns the next
 * mol_nsh[m] shells and mol_natom[m] nuclei of the concatenated arrays, and
 * its shell_ao values are relative to its own nbf = mol_nbf[m]. Matrices of
 * a batch plan (P, J, K, S, T, V) are packed: molecule m's nbf_m x nbf_m
 * row-major block starts at element sum_{m' < m} nbf_m'^2. All quartet
 * classes of all molecules are sorted together, so one dft_jk call (or one
 * graph launch) builds every molecule's J/K.
 *
 * Replaces running the reference's whole set-up and SCF once per molecule.
 */
int dft_plan_create_batch(int nmol, const int32_t* mol_nsh, const int32_t* mol_natom, const int32_t* mol_nbf,
                          const int32_t* shell_l, const int32_t* shell_nprim, const int32_t* shell_ao,
                          const double* shell_xyz, const double* prim_exp, const double* prim_coef,
                          const double* atom_xyz, const double* atom_charge, double schwarz_tau, double fp32_tau,
                          int flags, dft_plan* out);

int dft_plan_destroy(dft_plan plan);

typedef struct dft_graph_s* dft_graph;

/*
 * Capture one dft_jk (same arguments; the pointers are bound into the graph)
 * as an instantiated hipGraph. Launching the graph rebuilds J/K from the
 * current contents of P, Pa, Pb; update those buffers in place between
 * launches. No counterpart in the reference.
 */
int dft_jk_graph_create(dft_plan plan, const double* P, const double* Pa, const double* Pb, double* J, double* Ka,
                        double* Kb, int mode, dft_graph* out);
int dft_graph_launch(dft_graph graph, void* stream);
int dft_graph_destroy(dft_graph graph);

/*
 * Autotune the launch shape of every class-combination launch of the plan:
 * for the FP64 launch (whole list) and the FP32 launch (weak part) it times
 * the class's kernel (lane-per-quartet kernel, or the column kernel of the
 * heavy d classes) at workgroup sizes 64/128/256/512 (up to the class's
 * limit) and grid caps 512..8192 (median of reps runs each) and keeps the
 * fastest; then it times the FP64 part alone and keeps the FP32 launch for
 * DFT_MIXED only if FP64 part + FP32 part beat one FP64 launch over the whole
 * list by 3% (else that class runs all-FP64 in DFT_MIXED too); whole-plan
 * tuning (combo -1) keeps those choices only if whole mixed builds with them
 * beat whole mixed builds with FP32 in every class. For the heavy d
 * combinations it also times the matrix-core variant (FP64 MFMA products,
 * kernels_mfma.hip) of the FP64 part and keeps it where it is faster. report
 * (capacity report_cap ints) receives 6 ints per combination: FP64
 * {m, block, cap} with m = 1 where the matrix-core variant runs, FP32
 * {f, block, cap} with f = 0 (FP32 used) or 1 (FP64 instead). combo >= 0 tunes only that combination; -1 tunes
 * all of them and then chooses between running the class launches of a
 * build serially or concurrently (timing whole FP64 builds); -2 makes only
 * that last choice. Overwrites the plan's staging (J/K are not written).
 * No counterpart in the reference.
 */
int dft_plan_tune(dft_plan plan, const double* P, const double* Pa, const double* Pb, int reps, int combo,
                  int* report, int report_cap, void* stream);

/* Set launch shapes, the matrix-core choice and the FP32 choice (6 ints per
 * combination, as reported by dft_plan_tune). Every entry is checked before
 * any is applied: DFT_EINVAL leaves the plan as it was. */
int dft_plan_set_shapes(dft_plan plan, const int* shapes, int n);

/*
 * Plan statistics. counts receives DFT_STATS_COLS values for each class
 * combination c (pair classes ss=0, ps=1, pp=2, ds=3, dp=4, dd=5;
 * combination index pcb*(pcb+1)/2 + pck), at counts[DFT_STATS_COLS*c + k]:
 *   k=0 FP64 quartets, 1 FP32 quartets,
 *   2 primitive shell quartets of the FP64 part, 3 of the FP32 part (after
 *     primitive-pair screening and, for FP32, the FP32 primitive lists; the
 *     kernels' run-time primitive-quartet cut is not counted),
 *   4 and 5 the sums of K_a K_b K_c K_d over the same quartets (before any
 *     primitive screening),
 *   6 FP32 quartets on the second FP32 record level (Schwarz product below
 *     1e-3 of the FP32 threshold, in classes where that part is large),
 *   7 primitive shell quartets of the FP32 part with the FP64 records (what
 *     an all-FP64 build evaluates there).
 * ncombo is the capacity of counts in combinations; the number of
 * combinations is returned in *ncombo_out. Also reports the quartets before
 * screening and, in *staging_bytes, the bytes of device workspace one Fock
 * build uses (J/K staging slots, heavy-class integral
 * buffers): what dft_jk_static needs to fit in DFT_STATIC_SCRATCH_BYTES.
 */
#define DFT_STATS_COLS 8
int dft_plan_stats(dft_plan plan, int64_t* counts, int ncombo, int* ncombo_out, int64_t* n_unscreened,
                   int64_t* staging_bytes);

/*
 * Schwarz bound per input shell pair: Q[i*nsh + j] = Q[j*nsh + i] =
 * sqrt(max_{mu in i, nu in j} |(mu nu|mu nu)|). Q is a host array nsh*nsh.
 */
int dft_schwarz(dft_plan plan, double* Q);

/*
 * Direct J/K build: J = sum_cd (ab|cd) P_cd, K_a = sum_bd (ab|cd) Pa_bd and,
 * for a UHF plan, K_b from Pb. Fused ERI evaluation + digestion, then a
 * deterministic reduction. Overwrites J, Ka, Kb (no accumulation). The
 * class launches run concurrently on the caller's stream and the plan's
 * auxiliary streams, joined back into the caller's stream before the
 * reduction; calls on one plan must not overlap (they share the staging).
 * Primitive quartets whose primitive Schwarz bound is below 1e-14 are left
 * out (each integral changes by < 1.3e-11); in DFT_MIXED the FP32 quartets
 * use FP32 primitive lists cut to the same order of error relative to their
 * Schwarz bound (below 1e-11 absolute).
 *
 * Replaces fill_matrices(shells, P, P_a, P_b, J, K_a, K_b)
 * (shelled.py:770-877) together with the ERI store it reads
 * (calc_eri/process_eri, shelled.py:596-768), and the dense J/K of
 * main.py:787-789.
 *
 *  P, Pa, Pb   device nbf*nbf, symmetric (the eight-fold quartet symmetry
 *              is applied, so each density is read in one orientation only);
 *              Pb and Kb must be NULL for a non-UHF plan
 *  mode        DFT_FP64 or DFT_MIXED
 */
int dft_jk(dft_plan plan, const double* P, const double* Pa, const double* Pb, double* J, double* Ka,
           double* Kb, int mode, void* stream);

/*
 * dft_jk with the Fock-build workspace (J/K staging slots,
 * heavy-class integral buffers) in the library's static device scratch
 * instead of the plan's: allocates nothing, for callers that cannot call
 * hipMalloc (code under hipGraph capture, tight serving loops, harnesses
 * that forbid allocation). Same arguments, kernels and results as dft_jk
 * (bitwise equal J/K). The plan may be one built with
 * DFT_PLAN_STATIC_WORKSPACE (which then holds no workspace of its own).
 * Limit: the plan's workspace (dft_plan_stats *staging_bytes) must fit in
 * DFT_STATIC_SCRATCH_BYTES, else DFT_EINVAL; the call must be made with the
 * plan's device current. Every call on one GPU shares that GPU's scratch,
 * so calls that overlap in time on one GPU (other streams, other threads,
 * other plans) race: serialise them. Nothing in the scratch needs clearing
 * between calls (every slot is rewritten by each build).
 *
 * No counterpart in the reference: an allocation-free variant of
 * fill_matrices (shelled.py:770-877) as replaced by dft_jk.
 */
int dft_jk_static(dft_plan plan, const double* P, const double* Pa, const double* Pb, double* J, double* Ka,
                  double* Kb, int mode, void* stream);

/*
 * Analytic nuclear gradient of the (U)HF energy without the nuclear
 * repulsion term:
 *   grad = sum P dH/dR + sum Gamma d(ab|cd)/dR - sum W dS/dR
 * with Gamma the density pair of P (total), Pa, Pb (Pb NULL: Pb = Pa) and W
 * the energy-weighted density sum_s C_s,occ diag(eps_s) C_s,occ^T. grad is a
 * device array [3*natom]. shell_atom[i] is the nucleus index of input shell
 * i; the first call builds the gradient data of the plan (extended pair
 * records, slot map) and allocates, later calls only enqueue work and
 * reuse that mapping. s and p shells. Deterministic.
 *
 * No counterpart in the reference (it has no gradients); it is the backward
 * of the energy for torch.autograd.
 */
int dft_grad(dft_plan plan, const int32_t* shell_atom, const double* P, const double* Pa, const double* Pb,
             const double* W, double* grad, void* stream);

/*
 * Per-launch timing of the next nsteps dft_jk calls (0 disables). Creates
 * 2*nslot HIP events per step; no device memory. Slots: 2c = class
 * combination c, FP64 part; 2c+1 = its weak (FP32-eligible) part; then the
 * reduction and the symmetrisation. No counterpart in the reference (it has
 * no timing code, main.py:6).
 */
int dft_plan_profile(dft_plan plan, int nsteps);

/*
 * Sum over the recorded steps of each slot's duration in ms (synchronises
 * on the recorded events). ms_sum has capacity cap; *nslot and *nsteps
 * receive the slot count and the number of recorded steps.
 */
int dft_plan_profile_read(dft_plan plan, double* ms_sum, int cap, int* nslot, int* nsteps);

/*
 * Contracted, normalised ERI blocks for arbitrary shell quartets (input shell
 * indices). quartets is a host int32 array [4*nq]; out is a host array that
 * receives the blocks one after another, each laid out [n_a][n_b][n_c][n_d]
 * in the requested orientation. Synchronous.
 *
 * Replaces eri_kernel + the normalise/contract step of calc_eri
 * (shelled.py:452-532, 642-689).
 */
int dft_eri_quartets(dft_plan plan, const int32_t* quartets, int64_t nq, double* out);

/*
 * Overlap, kinetic and nuclear-attraction matrices (device nbf*nbf each).
 * Replaces overlap/T_raw/nuclear_kernel + contraction (shelled.py:962-1069).
 */
int dft_oneelec(dft_plan plan, double* S, double* T, double* V, void* stream);

/*
 * Values of every Cartesian basis function at npts points:
 * ao[mu * ld + g] = phi_mu(r_g) (device, nbf rows of stride ld >= npts),
 * grid_xyz device FP32 [npts][3] in bohr. Single-molecule plans.
 *
 * Replaces the AO evaluation of the density post-processing
 * (shelled_eval.py:98-113; flat variant main.py:905-912); the MO and density
 * contractions that follow there are plain GEMMs on these values.
 */
int dft_ao_grid(dft_plan plan, const float* grid_xyz, int64_t npts, double* ao, int64_t ld, void* stream);

/*
 * Primitive one-electron pass in the buffer layout of the reference's
 * nuclear_kernel: one entry q = a * nfb + b per primitive Cartesian pair.
 * Outputs V[q] = -(2 pi/p) sum_C Z_C sum_tuv E_t E_u E_v R_tuv(p, P - C),
 * the Hermite coefficients Ex/Ey/Ez[q*7 + t] (zero padded), p[q], P[3q+k]
 * and the lengths nx/ny/nz[q] = l_a + l_b + 1 per axis. cen_*_stride is 0
 * for one centre per side (shelled.py:350-371, per shell pair) or 3 for a
 * centre per function (main.py:392-409, flat variant). Device pointers;
 * l_a + l_b up to 12 (the reference's capacity).
 *
 * Replaces nuclear_kernel (shelled.py:350-450, main.py:392-488).
 */
int dft_nuclear_prim(int nfa, int nfb, const double* exp_a, const double* exp_b, const double* cen_a,
                     int cen_a_stride, const double* cen_b, int cen_b_stride, const int32_t* pow_a,
                     const int32_t* pow_b, int natom, const int32_t* atoms, const double* atom_xyz, double* V,
                     double* Ex, double* Ey, double* Ez, double* p, double* P, int32_t* nx, int32_t* ny, int32_t* nz,
                     void* stream);

/*
 * Unnormalised, uncontracted primitive Cartesian ERIs from two pair-record
 * sets in the reference layout: out[i * n_cd + j] = (i | j) for every bra
 * record i < n_ab and ket record j < n_cd. E arrays have stride e_stride (7
 * in the reference). Total momentum up to 12. Device pointers.
 *
 * Replaces eri_kernel (shelled.py:452-532; the flat main.py:490-563 variant
 * is the same call with the bra pointers offset by row_start).
 */
int dft_eri_prim(int64_t n_ab, const double* p_ab, const double* P_ab, const double* Ex_ab, const double* Ey_ab,
                 const double* Ez_ab, const int32_t* nx_ab, const int32_t* ny_ab, const int32_t* nz_ab, int64_t n_cd,
                 const double* p_cd, const double* P_cd, const double* Ex_cd, const double* Ey_cd,
                 const double* Ez_cd, const int32_t* nx_cd, const int32_t* ny_cd, const int32_t* nz_cd, int e_stride,
                 double* out, void* stream);

/*
 * Batched symmetric eigendecomposition for the SCF step: for each of the
 * batch matrices A_b (n x n, row-major, symmetric; overwritten), the
 * eigenvalues in ascending order in w[b*n ...] and the eigenvectors as the
 * columns of V_b (n x n, row-major). With init_v = 1, V is set to the
 * identity first; with init_v = 0, V must hold an orthonormal basis Q and A
 * the matrix in that basis (Q^T F Q), and the eigenvectors of F are
 * returned (Q times the rotations): a warm start from the previous
 * iteration's orbitals. sweeps (device, batch ints, or NULL) receives the
 * Jacobi sweeps each matrix took. Two-sided cyclic Jacobi, one workgroup
 * per matrix, n <= 128 (DFT_EUNSUPPORTED above). Only enqueues work (no
 * allocation, no synchronisation), so it can be captured into a hipGraph.
 * Device pointers.
 *
 * Replaces xp.linalg.eigh of the SCF loops (shelled.py:1210-1213,
 * main.py:796-799) and of S (shelled.py:1082, main.py:754).
 */
int dft_syevj(int n, int batch, double* A, double* w, double* V, int init_v, int32_t* sweeps, void* stream);

/*
 * dft_syevj with an absolute tolerance on the off-diagonal entries: pairs
 * with |a_pq| <= max(tol, n 2^-52 ||A||_F) are not rotated and the sweeps
 * stop once no entry is above that bound (tol = 0: dft_syevj). For an SCF
 * loop whose convergence test is far coarser than machine precision (its
 * orbitals then carry errors of about tol / gap, its energies of about
 * (tol / gap)^2 at convergence); DFT_EINVAL for tol < 0 or NaN.
 *
 * Replaces the same eigendecompositions as dft_syevj (shelled.py:1210-1213,
 * main.py:796-799).
 */
int dft_syevj_tol(int n, int batch, double* A, double* w, double* V, int init_v, int32_t* sweeps, double tol,
                  void* stream);

/*
 * Bytes of device workspace dft_syevj_ws takes per matrix of order n (0 for
 * n outside 1..128): a rotation log of 40 sweeps of (n + n mod 2 - 1)
 * rounds of (n + n mod 2) / 2 rotations, 16 bytes each, plus a 1 KiB header;
 * 4.0 MB at n = 112.
 */
int64_t dft_syevj_ws_bytes(int n);

/*
 * dft_syevj_tol with a device workspace (work, work_bytes >= batch x
 * dft_syevj_ws_bytes(n), or NULL for dft_syevj_tol itself). The matrix
 * workgroup then keeps no eigenvectors: it records every round's rotations
 * in the workspace, and a second kernel applies them to the rows of V, one
 * wave per row (the same rotations in the same order and arithmetic). For
 * small batches (the SCF step), where the eigenvector update shares the
 * matrix workgroup's LDS bandwidth. Same results and error behaviour as
 * dft_syevj_tol; DFT_EINVAL for a workspace below the size. Still only
 * enqueues work (two kernels), so it can be captured into a hipGraph.
 *
 * Replaces the same eigendecompositions as dft_syevj (shelled.py:1210-1213,
 * main.py:796-799).
 */
int dft_syevj_ws(int n, int batch, double* A, double* w, double* V, int init_v, int32_t* sweeps, double tol,
                 void* work, int64_t work_bytes, void* stream);

/*
 * One damped + DIIS SCF iteration on the device, closed shell (the loop of
 * shelled.py:1164-1260 with P_a = P_b), split in two calls around the
 * orbital step so that a whole iteration (dft_jk, these two, dft_syevj and
 * GEMMs) can be captured once and replayed as a hipGraph: the loop's control
 * state lives in the device array state[DFT_SCF_STATE_LEN] (indices
 * DFT_SCF_ST_*), read by the host once per iteration. All matrices are
 * device n x n row-major; hist_F, hist_R are rings of `history`
 * (<= DFT_SCF_HMAX) matrices, G the history x history Gram matrix, all
 * zero-initialised by the caller; state[DFT_SCF_ST_RATE, _ENN, _ETOL,
 * _PTOL, _DIIS_TH] are set by the caller, the other entries start at 0
 * (PREV_DE and PREV_DP at +inf); entries from DFT_SCF_ST_COEF on hold the
 * DIIS coefficients and per-workgroup partial sums (scratch of the calls).
 *
 * dft_scf_diis_stage: stores F and R = F Pa S - S Pa F in the ring slot
 * count mod history, updates G's row of that slot, solves the DIIS
 * equations (B c = 1 without a Lagrange row, c normalised to sum 1,
 * shelled.py:1196-1203) on the used slots, writes the iteration's Fock
 * matrix Fu (the extrapolation once DIIS is on, else F) and
 * state[DFT_SCF_ST_ESUM] = sum P.H + 2 sum Pa.Fu.
 * dft_scf_damp_control: from the aufbau density Pauf of Fu, the damped
 * density rate Pauf + (1 - rate) Pa (UHF_density, shelled.py:561-574),
 * E = ESUM / 2 + E_NN, dE, dP = max |change|, the strike / rate / DIIS
 * switch rules (shelled.py:1250-1257) and the convergence flag; unless
 * converged, Pa and P = 2 Pa take the damped density (converged: Pa stays
 * the density of the reference's final rebuild, shelled.py:1261-1285).
 *
 * Replace the per-iteration host code of shelled.py:1164-1260 around
 * fill_matrices (no counterpart as device code).
 */
#define DFT_SCF_HMAX 16
#define DFT_SCF_STATE_LEN 1152
#define DFT_SCF_ST_ESUM 0
#define DFT_SCF_ST_ENN 1
#define DFT_SCF_ST_E 2
#define DFT_SCF_ST_HAS_E 3
#define DFT_SCF_ST_DE 4
#define DFT_SCF_ST_DP 5
#define DFT_SCF_ST_PREV_DE 6
#define DFT_SCF_ST_PREV_DP 7
#define DFT_SCF_ST_RATE 8
#define DFT_SCF_ST_STRIKES 9
#define DFT_SCF_ST_USE_DIIS 10
#define DFT_SCF_ST_COUNT 11
#define DFT_SCF_ST_CONV 12
#define DFT_SCF_ST_ETOL 13
#define DFT_SCF_ST_PTOL 14
#define DFT_SCF_ST_DIIS_TH 15
#define DFT_SCF_ST_ENEW 16
#define DFT_SCF_ST_SLOT 17
#define DFT_SCF_ST_COEF 32
int dft_scf_diis_stage(int n, int history, const double* F, const double* R, double* hist_F, double* hist_R,
                       double* G, const double* P, const double* H, const double* Pa, double* state, double* Fu,
                       void* stream);
int dft_scf_damp_control(int n, double* Pa, const double* Pauf, double* P, double* state, void* stream);

/*
 * The dense products around those two calls, for orders up to 128 (one
 * 16 x 16 output tile per workgroup, each output one FMA chain over k in
 * order; DFT_EUNSUPPORTED above 128). Device pointers, row-major; they only
 * enqueue work, so a whole iteration can be captured.
 *
 * dft_scf_gemm: C (m x ncol, ldc) = op(A) op(B) over k, with op(A) = A
 * (m x k, lda) or A^T (trans_a: A is k x m), op(B) = B (k x ncol, ldb) or
 * B^T (trans_b: B is ncol x k). With copy_src (and copy_dst) non-NULL, also
 * copies the m x ncol block copy_src to copy_dst (leading dimension ldc).
 * Replaces the matrix products of the SCF step (P_a S, X Q, F X Q, the
 * orbitals X C and the aufbau density C_occ C_occ^T; shelled.py:1191-1193,
 * 1210-1217).
 *
 * dft_scf_fock_residual: F = (H + J) - K and the DIIS residual
 * R = F PS - PS^T F for PS = P_a S (shelled.py:1179-1180, 1191-1193), each
 * n x n.
 *
 * dft_scf_orbital_matrix: A = (At + At^T) / 2 for At = XQ^T T (T = F XQ:
 * the Fock matrix in the previous orbitals' basis, shelled.py:1210-1211) and
 * V = Q (the eigensolver's warm start), each n x n.
 */
int dft_scf_gemm(int m, int ncol, int k, const double* A, int lda, int trans_a, const double* B, int ldb,
                 int trans_b, double* C, int ldc, const double* copy_src, double* copy_dst, void* stream);
int dft_scf_fock_residual(int n, const double* H, const double* J, const double* K, const double* PS, double* F,
                          double* R, void* stream);
int dft_scf_orbital_matrix(int n, const double* XQ, const double* T, double* A, const double* Q, double* V,
                           void* stream);

/*
 * Boys function test hook: out[i*(mmax+1)+m] = F_m(T[i]), m = 0..mmax, with
 * the tabulated evaluator the ERI kernels use. T and out are device arrays.
 * fp32 selects the FP32 recursion (FP64 seed). Replaces boys/taylor
 * (shelled.py:174-211).
 */
int dft_boys(const double* T, int n, int mmax, double* out, int fp32, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DFT_HIP_H */
