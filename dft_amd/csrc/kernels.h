// Host-visible launch entry points of the device kernels (implemented in the
// kernels_*.hip translation units). All pointers are device pointers.
//
// Matrices (densities, J, K, S, T, V) are packed per molecule: molecule m's
// nbf_m x nbf_m row-major block starts at element mbase_m. A single-molecule
// plan is the case of one block at offset 0. Every shell carries its
// molecule's (mbase, nbf) in a packed int4 "shell_dm" = {ao, nbf, mbase lo,
// mbase hi}.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dft {

// number of pair classes / class combinations for shells up to l = 3 (s, p,
// d, f; combinations with an f shell run the run-time-momentum kernels of
// kernels_generic.hip, the others compile-time ones)
constexpr int LMAX = 3;
constexpr int NPC = (LMAX + 1) * (LMAX + 2) / 2;   // pair classes: ss, ps, pp, ds, dp, dd, fs, fp, fd, ff
constexpr int NCOMBO = NPC * (NPC + 1) / 2;        // (bra class >= ket class)

// lanes of a wavefront, i.e. quartets per wave of a lane-kernel launch; the
// lane-kernel quartet lists are padded to whole waves per molecule and part
constexpr int WAVE_QUARTETS = 64;

inline int pc_index(int la, int lb) { return la * (la + 1) / 2 + lb; }
inline void pc_moments(int pc, int& la, int& lb) {
  la = 0;
  while ((la + 1) * (la + 2) / 2 <= pc) ++la;
  lb = pc - la * (la + 1) / 2;
}
inline int combo_index(int pcb, int pck) { return pcb * (pcb + 1) / 2 + pck; }
inline int combo_pcb(int combo) {
  int b = 0;
  while ((b + 1) * (b + 2) / 2 <= combo) ++b;
  return b;
}
inline int combo_pck(int combo) { return combo - combo_pcb(combo) * (combo_pcb(combo) + 1) / 2; }

struct ShellDev {      // one contracted (segmented) shell, internal order
  int l, nprim, prim_off, ao;
  double x, y, z;
  int atom_begin, atom_end;  // nuclei of the shell's molecule (for V)
  int nbf, mol;
  long long mbase;           // packed-matrix offset of the molecule
};

struct Target {        // one ordered shell pair (X, Y) of a half matrix H (J = H + H^T)
  long long base;      // pad * nelem + offset of element (X row 0, Y col 0) in its matrix
  int nrow, ncol, stride;
  int pad;             // the matrix: 0 = J, 1 + s = K of density s
};

// pair pre-pass: one lane per primitive pair; writes FP64 and / or FP32
// records (either output may be null), with Schwarz factors sq (or 0 if null)
hipError_t launch_pair_records(const int4* pairs_prim /* sh_a, sh_b, i_a, i_b */, int nprim_pairs,
                               const int* rec_off /* per primitive pair, in reals */,
                               const ShellDev* shells, const double* exps, const double* coefs,
                               const double* sq, double* rec64, float* rec32, hipStream_t s);

// true if the lane-per-quartet kernel is instantiated for this combination
// (s/p classes and the light d classes); the others use the column kernel
// (launch_eri_chunk) and its digest
bool lane_kernel_combo(int combo);

// Pair records of one lane-kernel launch. The bra pairs of a wave differ
// from lane to lane and are read from the interleaved block of their class
// (element f of primitive pair i of class-local pair u at
// bra[(i * RS + f) * sb + u]; pair id = fb + u), so those loads coalesce;
// the ket pair is shared by the wave (ket-major quartet order) and is read
// from the plain records (ket + pair.z), a few cache lines per pair.
struct RecView {
  const void* bra;
  int sb, fb;
  const void* ket;
  // quartets [first, first + n) of the combination's list of ntot, each
  // evaluated by splb * splk lanes (primitive split)
  int first, n, ntot, splb, splk;
  // Lane table of a split combination (nullptr: one lane per quartet and
  // chunk as above). Entry i = (quartet, first bra primitive, bra primitive
  // stride, idle flag): the entry takes bra primitive pairs first, first +
  // stride, ...; the lanes of one quartet are adjacent in a wave, so
  // the in-wave run sums add its bra chunks before anything is stored. With
  // a table, first / n / ntot count table entries, and every entry runs
  // splk times (ket chunk ck = chunk-major index, one slot row each).
  const int4* lanes;
};

// Copies plain pair records (one record of rs reals per primitive pair at
// src_off) into the interleaved layout: map[r] = {src_off, dst_off, rs, s}
// with element f going to dst_off + f * s.
hipError_t launch_interleave_records(const int4* map, int nrec, const double* r64, const float* r32, double* t64,
                                     float* t32, hipStream_t s);

// Fused ERI + J/K digest for one class combination (lane per quartet or per
// quartet primitive chunk); slots holds 2 + 4*nk staging offsets (doubles)
// per quartet (-1 where the lane's block is summed into a later lane's; the
// primitive chunk ch of a quartet adds ch times the block size); segmask one
// word per (wave of the list, block kind) with bit l set where lane l starts
// a run of lanes with the same target. With a lane table (rv.lanes) slots
// and segmask are per table entry and per wave of the table, and only the
// ket chunk adds to the slot; block is the workgroup size
// (64 ... 512, at most lane_block_max), grid_cap the maximum number of
// workgroups. Quartets with a negative bra pair are padding (no work, no
// stores).
// largest lane-kernel workgroup of a class combination (0 if none)
int lane_block_max(int combo);

hipError_t launch_eri_jk(int combo, bool fp32, const int2* quartets, const long long* slots,
                         const unsigned long long* segmask, const int4* pairs, const RecView& rv,
                         const int4* shell_dm,
                         const double* DJ, const double* DK0, const double* DK1, int nk,
                         double* stage, const double* boys_tab, int block, int grid_cap, hipStream_t s);

// Heavy combinations (d with p/d; no lane kernel): integrals of quartets
// [first, first + n), one lane per (bra primitive chunk sp < splb, quartet,
// ket component pair), chunk sp's partial integrals stored at
// out[sp * ss + x * xs + q * qs] (x = abcd row-major); and the J/K blocks of
// the whole list of n quartets from integrals stored with xs = 1, qs = NI,
// summed over the splb chunks.
hipError_t launch_eri_chunk(int combo, bool fp32, const int2* quartets, int first, int n, int splb,
                            const int4* pairs, const RecView& rv, double* out, long long xs, long long qs,
                            long long ss, const double* boys_tab, int block, int grid_cap, bool prim_cut,
                            hipStream_t s);
hipError_t launch_jk_digest(int combo, const int2* quartets, const long long* slots, int n, int splb, long long ss,
                            const int4* pairs, const int4* shell_dm, const double* DJ, const double* DK0,
                            const double* DK1, int nk, const double* ibuf, double* stage, hipStream_t s);

// Heavy combinations on the matrix cores (FP64 MFMA products for the ket and
// bra steps, kernels_mfma.hip): the same integrals and output layout as
// launch_eri_chunk, one wave per (quartet, bra primitive chunk).
bool mfma_combo(int combo);
hipError_t launch_eri_mfma(int combo, const int2* q, int first, int n, int splb, const int4* pairs, const double* rec,
                           double* out, long long xs, long long qs, long long ss, const double* gtab, int cap,
                           bool cut, hipStream_t s);

// Class combinations with an f shell (kernels_generic.hip): one wave per
// quartet, momenta at run time; integrals to out[x * xs + q * qs] (q the
// list index, quartets [first, first + n)), plain FP64 records; FP64 only.
// cut: leave out primitive quartets below the J/K kernels' primitive cut
// (needs the records' Schwarz factors). The J/K blocks of a whole list from
// integrals stored with xs = 1, qs = ni, slots quartet-major.
bool generic_combo(int combo);
size_t generic_lds_bytes(int la, int lb, int lc, int ld);
hipError_t launch_eri_generic(int la, int lb, int lc, int ld, const int2* quartets, int first, int n,
                              const int4* pairs, const double* rec, double* out, long long xs, long long qs,
                              const double* boys_tab, int grid_cap, bool cut, hipStream_t s);
hipError_t launch_jk_generic(int la, int lb, int lc, int ld, const int2* quartets, const long long* slots, int n,
                             const int4* pairs, const int4* shell_dm, const double* DJ, const double* DK0,
                             const double* DK1, int nk, const double* ibuf, double* stage, hipStream_t s);

// raw contracted blocks (internal orientation) for one class combination:
// out + q * NA*NB*NC*ND
hipError_t launch_eri_raw(int combo, const int2* quartets, long long nq, const int4* pairs,
                          const double* rec, double* out, const double* boys_tab, hipStream_t s);

// Deterministic reduction and symmetrisation: one wave per unordered
// shell pair tpairs[i] = (target XY, target YX) of one matrix (YX = XY on
// the diagonal); the blocks of target t are contiguous:
// stage[tgt_base[t] + k * nrow * ncol + e], k < tgt_count[t]. Writes both
// orientations of out = H + H^T of the target's matrix (Target.pad: 0 = J
// to out0, 1 = K_a to out1, 2 = K_b to out2; Target.base - pad * nelem is
// the element offset in that matrix).
hipError_t launch_reduce_sym(int npair, const int2* tpairs, const Target* targets, const long long* tgt_base,
                             const int* tgt_count, const double* stage, long long stage_len /* doubles */,
                             long long nelem, double* out0, double* out1, double* out2, hipStream_t s);

// AO values on a grid: out[ao * ld + g], g < npts (grid: npts x 3 FP32)
hipError_t launch_ao_grid(int nsh, const ShellDev* shells, const double* exps, const double* coefs,
                          const float* grid, long long npts, double* out, long long ld, hipStream_t s);

// S, T, V of the shell pairs sp[0..npairs) of one class (la >= lb; the
// first shell of each pair has momentum la), one 64-lane workgroup per
// shell pair, packed outputs (kernels_onee.hip)
hipError_t launch_one_electron(int la, int lb, const int2* shell_pairs, int npairs, const ShellDev* shells,
                               const double* exps, const double* coefs, const double* atom_xyzq, double* S,
                               double* T, double* V, const double* boys_tab, hipStream_t s);

// largest shell momentum of the analytic gradient kernels
constexpr int GRAD_LMAX = 1;

// analytic gradients (s/p, one molecule): extended pair records, derivative
// ERIs contracted with the density pair (bra-centre derivatives; call again
// with bra and ket exchanged for the ket centres), 1-e derivatives and the
// per-atom reduction. Every contribution is 3 doubles at its own slot.
hipError_t launch_grad_records(const int4* pairs_prim, int n, const int* rec_off, const ShellDev* shells,
                               const double* exps, const double* coefs, double* out, hipStream_t s);
hipError_t launch_eri_grad(int la, int lb, int lc, int ld, const int2* quartets, const long long* slots, long long nq,
                           const int4* gpairs, const double* grec, const int* shell_ao, const double* DJ,
                           const double* DK0, const double* DK1, int nbf, double* stage, const double* boys_tab,
                           hipStream_t s);
// 1-e derivatives of the shell pairs sp[pidx[0..npairs)] of one class
// (la >= lb), one 64-lane workgroup per shell pair; slots: three per
// (shell pair, nucleus) (kernels_onee.hip)
hipError_t launch_grad1e(int la, int lb, const int2* shell_pairs, const int* pidx, int npairs, const ShellDev* shells,
                         const double* exps, const double* coefs, int natom, const double* atoms, const double* P,
                         const double* W, int nbf, const long long* slots, double* stage, const double* boys_tab,
                         hipStream_t s);
// GRAD_CHUNK: slots per stage-1 chunk of the per-atom gradient sums
constexpr int GRAD_CHUNK = 1024;
hipError_t launch_grad_reduce(int natom, int nchunk, const int4* chunks, const int* cfirst, const long long* base,
                              const double* stage, double* part, double* grad, hipStream_t s);

// reference-layout primitive kernels (nuclear_kernel / eri_kernel buffers)
hipError_t launch_nuclear_prim(int nfa, int nfb, const double* ea, const double* eb, const double* ca, int sa,
                               const double* cb, int sb, const int* pa, const int* pb, int natom, const int* Z,
                               const double* xyz, double* V, double* Ex, double* Ey, double* Ez, double* p,
                               double* P, int* nx, int* ny, int* nz, const double* boys_tab, hipStream_t s);
hipError_t launch_eri_prim(long long nab, const double* pab, const double* Pab, const double* Exab,
                           const double* Eyab, const double* Ezab, const int* nxab, const int* nyab,
                           const int* nzab, long long ncd, const double* pcd, const double* Pcd, const double* Excd,
                           const double* Eycd, const double* Ezcd, const int* nxcd, const int* nycd,
                           const int* nzcd, int e_stride, double* out, const double* boys_tab, hipStream_t s);

// Batched symmetric eigensolver (two-sided cyclic Jacobi, one workgroup per
// matrix, order n <= JAC_NMAX): A (batch x n x n, row-major, destroyed), w
// (batch x n, ascending), V (batch x n x n, eigenvectors as columns; with
// init_v = 0 it holds an orthonormal starting basis on entry and the result
// is that basis times the rotations). With a rotation log (jlog, at least
// syevj_log_bytes(n) bytes per matrix, jlog_stride apart) the matrix
// workgroup keeps no eigenvectors: it logs its rotations and a second kernel
// applies them to the rows of V in parallel.
constexpr int JAC_NMAX = 128;
long long syevj_log_bytes(int n);
hipError_t launch_syevj(int n, int batch, double* A, long long sa, double* w, long long sw, double* V,
                        long long sv, int init_v, int* sweeps, double tol, char* jlog, long long jlog_stride,
                        hipStream_t s);

// device SCF iteration (state layout: the DFT_SCF_ST_* indices of dft_hip.h)
constexpr int SCF_HMAX = 16;
enum ScfState {
  SCF_ST_ESUM = 0, SCF_ST_ENN = 1, SCF_ST_E = 2, SCF_ST_HAS_E = 3, SCF_ST_DE = 4, SCF_ST_DP = 5,
  SCF_ST_PREV_DE = 6, SCF_ST_PREV_DP = 7, SCF_ST_RATE = 8, SCF_ST_STRIKES = 9, SCF_ST_USE_DIIS = 10,
  SCF_ST_COUNT = 11, SCF_ST_CONV = 12, SCF_ST_ETOL = 13, SCF_ST_PTOL = 14, SCF_ST_DIIS_TH = 15,
  SCF_ST_ENEW = 16, SCF_ST_SLOT = 17,
  // DIIS coefficients (SCF_HMAX), then per-workgroup partial sums of the
  // element-wise passes
  SCF_ST_COEF = 32, SCF_ST_PART = 48, SCF_STATE_LEN = 1152
};
hipError_t launch_scf_diis_stage(int n, int h, const double* F, const double* R, double* histF, double* histR,
                                 double* G, const double* P, const double* H, const double* Pa, double* st,
                                 double* Fu, hipStream_t s);
hipError_t launch_scf_damp_control(int n, double* Pa, const double* Pauf, double* P, double* st, hipStream_t s);
// dense products of the iteration (orders <= JAC_NMAX, 16 x 16 tiles)
hipError_t launch_scf_gemm(int m, int nc, int kd, const double* A, int lda, int ta, const double* B, int ldb, int tb,
                           double* C, int ldc, const double* cp_src, double* cp_dst, hipStream_t s);
hipError_t launch_scf_fock_residual(int n, const double* H, const double* J, const double* K, const double* PS,
                                    double* F, double* R, hipStream_t s);
hipError_t launch_scf_orbital_matrix(int n, const double* XQ, const double* T, double* A, const double* Q, double* V,
                                     hipStream_t s);

// F_m(T) test hook
hipError_t launch_boys(const double* T, int n, int mmax, double* out, int fp32,
                       const double* boys_tab, hipStream_t s);

}  // namespace dft
