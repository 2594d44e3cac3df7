// C ABI of libdft_hip.so: plan construction (pair pre-pass, Schwarz bounds,
// class-sorted screened quartet lists, reduction map) for one molecule or a
// packed batch of molecules, the per-iteration entry points, hipGraph
// capture of the Fock build and the launch-shape autotuner. Declarations and
// conventions: include/dft_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dft_hip.h"
#include "kernels.h"
#include "md_device.h"

using namespace dft;

namespace {

// Primitive split of the lane-kernel combinations: a class with fewer than
// SPLIT_LANES quartets gets about SPLIT_LANES lanes, but every lane of a
// split quartet repeats the J/K digest and the run sums of its blocks, which
// for a (pp|pp) quartet cost about as much as three primitive quartets, so
// no split leaves a lane fewer than SPLIT_MIN_PRIM primitive quartets
// (K_a K_b K_c K_d, class average). (H2O)16 STO-3G (pp|pp), 81 primitive
// quartets per quartet: 27 lanes per quartet (3 each) 42 us, 45 lanes 74 us;
// C12H26 s/p/d (pp|pp), 12.8 k quartets of up to 1296: two lanes per quartet
// 1.26 ms. A pair split of (ps|ps) and (pp|ss) on (H2O)16 (126 k and 79 k
// quartets) shortens their serial launches (63.7 -> 55.4 us, 59.1 -> 53.9
// us) but lengthens the concurrent build (0.321 -> 0.330 ms), where the
// large classes already fill the chip and only the added digest work
// remains, so classes at or above SPLIT_LANES quartets stay unsplit.
constexpr long long SPLIT_LANES = 65536;
constexpr long long SPLIT_MIN_PRIM = 3;
// the same for the column kernel of the heavy classes (longer lanes)
constexpr long long SPLIT_LANES_HEAVY = 262144;
// FP32 records drop primitive pairs whose summed Schwarz bounds stay within
// this fraction of the shell pair's (FP32 resolution is 6e-8)
constexpr double FP32_PRIM_REL = 1e-7;
// FP32 quartets with Q_ab Q_cd below FP32_LEVEL2 * tau32 use a second record
// set cut at FP32_PRIM_REL / FP32_LEVEL2 (the same absolute bound)
constexpr double FP32_LEVEL2 = 1e-3;
constexpr long long FP32_LEVEL2_MIN = 1 << 20;  // quartets
// auxiliary streams per plan for concurrent class launches
constexpr int JK_AUX = 3;
// screened quartets below which a plan's class launches run concurrently by
// default (larger builds fill the chip with each launch and run faster serially)
constexpr long long CONCURRENT_QUARTETS = 8000000;

}  // namespace

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return fail(DFT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------- Boys
// F_m(T_k) = e^{-T} sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1)), a series of
// positive terms, summed in extended precision.
std::vector<double> boys_table_host() {
  std::vector<double> tab((size_t)BOYS_NG * BOYS_MT);
  for (int k = 0; k < BOYS_NG; ++k) {
    const long double T = (long double)k / 10.0L;
    const long double e = expl(-T);
    for (int m = 0; m < BOYS_MT; ++m) {
      long double term = 1.0L / (2 * m + 1), sum = term;
      for (int i = 1; i < 2000; ++i) {
        term *= 2.0L * T / (2 * m + 2 * i + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
      }
      tab[(size_t)k * BOYS_MT + m] = (double)(e * sum);
    }
  }
  return tab;
}

std::mutex g_boys_mu;
std::map<int, double*> g_boys_dev;

int boys_table_device(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_boys_mu);
  auto it = g_boys_dev.find(dev);
  if (it != g_boys_dev.end()) {
    *out = it->second;
    return DFT_OK;
  }
  static const std::vector<double> host = boys_table_host();
  double* d = nullptr;
  HIP_TRY(hipMalloc(&d, host.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(d, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  g_boys_dev[dev] = d;
  *out = d;
  return DFT_OK;
}

// Radial part of the primitive normalisation of shelled.py:95-101:
// (2a/pi)^{3/4} (4a)^{l/2}; the angular part 1/sqrt(prod (2l_d-1)!!) is
// applied per axis in the pair records.
double radial_norm(double a, int l) { return std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * l); }

template <typename T>
int upload(T** d, const std::vector<T>& h) {
  *d = nullptr;
  if (h.empty()) return DFT_OK;
  HIP_TRY(hipMalloc(d, h.size() * sizeof(T)));
  HIP_TRY(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return DFT_OK;
}

int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

// workspace regions start on 256-byte boundaries
long long ws_align(long long n) { return (n + 31) / 32 * 32; }

// Host threads for the plan build's loops over quartets: the hardware's,
// at most 16 (a GPU's share of a node's cores) and at most OMP_NUM_THREADS
// where that is set.
unsigned plan_threads() {
  unsigned n = std::max(1u, std::thread::hardware_concurrency());
  n = std::min(n, 16u);
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) n = std::min(n, (unsigned)v);
  }
  return n;
}

// fn(chunk, begin, end) over [0, n) in at most plan_threads() contiguous
// chunks, one thread each (small ranges run on the calling thread).
template <class F>
void parallel_chunks(size_t n, size_t min_per_chunk, F&& fn) {
  const size_t nt = std::min<size_t>(plan_threads(), std::max<size_t>(1, n / std::max<size_t>(1, min_per_chunk)));
  if (nt <= 1) {
    fn(0, 0, n);
    return;
  }
  const size_t chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b < e) th.emplace_back([&fn, t, b, e] { fn(t, b, e); });
  }
  for (auto& x : th) x.join();
}

size_t parallel_chunk_count(size_t n, size_t min_per_chunk) {
  return std::min<size_t>(plan_threads(), std::max<size_t>(1, n / std::max<size_t>(1, min_per_chunk)));
}

constexpr double PRIM_CUT = 39.1439465808987;  // -ln(1e-17)

// launch shape of one class-combination launch (chosen by dft_plan_tune)
struct Shape {
  int wave = 0;         // reserved (0)
  int block = 256;      // lane kernel workgroup size
  int cap = 256 * 16;   // maximum workgroups (persistent grid-stride loop)
};

struct Combo {
  int pcb = 0, pck = 0;
  std::vector<int2> q;         // (bra pair, ket pair) plan pair ids: FP64 part, then weak part
  // lane-kernel lists are padded: every molecule's FP64 part and FP32
  // level parts start on a wave (WAVE_QUARTETS) boundary, the gaps filled
  // with dummy quartets (bra pair -1), so the quartets a wave holds are the
  // same in every launch split (FP64 / mixed, single / batch plan)
  long long n64 = 0;           // size of the FP64 part (padded)
  long long n32a = 0;          // size of the first FP32 level (padded; the rest is level 2)
  long long r64 = 0, r32a = 0, r32b = 0;  // quartets without the padding
  // per (wave, block kind): bit l set where lane l starts a run of lanes
  // whose block goes to the same target (the run's last lane stores its sum)
  unsigned long long* d_segmask = nullptr;
  int2* d_q = nullptr;
  long long* d_slots = nullptr;  // per quartet 2 + 4 nk staging offsets
  long long prim64 = 0, prim32 = 0;      // primitive shell quartets evaluated
  long long prim64w = 0;                 // FP64 primitive quartets of the FP32-eligible part
  long long kprod64 = 0, kprod32 = 0;    // sum of K_a K_b K_c K_d (before primitive screening)
  Shape shape64, shape32;        // for the FP64 launch and the FP32 (weak part) launch
  // mixed mode runs the weak part in FP32 (the tuner clears this where one
  // FP64 launch over the whole list is faster than the two launches)
  bool use32 = true;
  // heavy combinations: the FP64 part on the matrix cores (chosen by the
  // tuner where it is faster than the column kernel)
  bool mfma = false;
  // primitive split of the lane kernel: every quartet is evaluated by
  // splb * splk lanes, lane (cb, ck) taking bra chunk cb and ket chunk ck of
  // the primitive pairs and writing its own partial J/K blocks; chunk-major
  // slot order (slot row = (cb * splk + ck) * q.size() + quartet)
  int splb = 1, splk = 1;
  // Lane table of a split lane-kernel combination (empty: the chunk-major
  // lanes above): entry (quartet, first bra primitive, bra primitive stride,
  // idle), the bra chunks of a quartet in neighbouring entries of one wave
  // (so the in-wave run sums add them) and splk ket chunks of each entry;
  // waves end at molecule and part boundaries. lpos: table positions of
  // the quartet positions 0, n64, n64 + n32a and q.size().
  std::vector<int4> lanes;
  int4* d_lanes = nullptr;
  long long lpos[4] = {0, 0, 0, 0};
  // heavy (column-kernel) combinations: integrals of the whole list,
  // q.size() x NI x splb doubles at this offset of the workspace's integral
  // region, read by the digest kernel
  long long ibuf_off = 0;
};

// Device scratch of one Fock build: the J/K staging slots and the heavy
// classes' integral buffers. A plan owns
// one (dft_jk); dft_jk_static points the same regions into the library's
// static scratch array.
struct Workspace {
  double* stage = nullptr;
  double* ibuf = nullptr;
};

}  // namespace

struct dft_plan_s {
  int nsh = 0, natom = 0, nk = 1, nmol = 1;
  int device = 0;                // HIP device the plan was built on
  bool split = true;             // primitive split of small lane-kernel classes
  bool mfma = false;             // heavy combinations' FP64 part on the matrix cores
  // class launches of a build on several streams: helps builds whose class
  // launches cannot fill the chip one at a time; large builds run faster
  // serially (set by build size, then by dft_plan_tune)
  bool concurrent = true;
  long long nelem = 0;           // packed matrix elements (sum of nbf_m^2)
  int max_nbf = 0;
  double tau = 0, tau32 = 0;
  std::vector<int> mol_nbf;
  std::vector<long long> mol_base;
  std::vector<ShellDev> sh;      // internal (l-sorted) order
  std::vector<int> in2int;       // input shell -> internal shell
  std::vector<double> exps, coefs;
  std::vector<int4> pairs;       // (sh_a, sh_b, rec_off, nprim_pairs), class-sorted
  // the same for the two FP32 record levels (looser primitive cuts; level 1
  // serves quartets with Q_ab Q_cd < FP32_LEVEL2 * tau32)
  std::vector<int4> pairs32[2];
  std::vector<int2> prim_list;   // surviving primitive pairs (ia, ib) of all pairs
  std::vector<int> pair_prim_begin;  // first prim_list entry of each pair
  std::vector<int> pair_pc;      // pair class per pair
  std::map<long long, int> pair_of;  // internal (i,j), i >= j -> pair id
  std::vector<double> Q;         // Schwarz per pair
  std::vector<Combo> combos;
  long long n_unscreened = 0;
  long long stage_len = 0;       // doubles
  int ntarget = 0;
  // device
  ShellDev* d_sh = nullptr;
  double* d_exps = nullptr;
  double* d_coefs = nullptr;
  int4* d_pairs = nullptr;
  int4* d_pairs32[2] = {};
  int4* d_shell_dm = nullptr;
  int* d_shell_ao = nullptr;
  double* d_rec64 = nullptr;
  float* d_rec32[2] = {};
  // interleaved copies of the records, one block per pair class (lane kernels)
  double* d_recT64 = nullptr;
  float* d_recT32[2] = {};
  long long cls_base[NPC] = {}, cls_base32[2][NPC] = {};
  int cls_first[NPC] = {}, cls_n[NPC] = {};
  // the plan's own workspace (one allocation; not made for a plan built with
  // DFT_PLAN_STATIC_WORKSPACE) and the sizes of its regions in doubles
  double* d_ws = nullptr;
  Workspace ws;
  long long ibuf_len = 0;
  bool own_ws = true;
  Target* d_tg = nullptr;
  long long* d_tbase = nullptr;
  int* d_tcount = nullptr;
  int2* d_tpairs = nullptr;  // (target XY, target YX) per unordered shell pair and matrix
  int ntpair = 0;
  long long* d_mol_base = nullptr;
  int* d_mol_nbf = nullptr;
  double* d_atoms = nullptr;     // x, y, z, Z
  int2* d_shpairs = nullptr;     // shell pairs (later shell first), sorted by class
  int nshpairs = 0;
  int shp_first[NPC] = {}, shp_n[NPC] = {};  // class ranges of d_shpairs
  double* d_boys = nullptr;
  // optional per-launch timing: slot 2c = combo c FP64 launch, 2c+1 = its
  // FP32 launch, then reduce, symmetrize; one begin/end event pair per slot and step
  int prof_steps = 0, prof_used = 0;
  std::vector<hipEvent_t> prof_ev;
  // class launches of one Fock build are spread over the caller's stream and
  // JK_AUX more streams (fork/join by events; all default priority)
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  // gradient data, built by the first dft_grad call (single molecule)
  bool grad_ready = false;
  std::vector<int> grad_atom;     // internal shell -> atom
  double* d_grec = nullptr;
  int4* d_gpairs = nullptr;
  std::vector<int2*> d_gq, d_gq_swap;  // per combination: quartets (no padding), bra and ket exchanged
  std::vector<long long> gq_n;
  std::vector<long long*> d_gslots_ab, d_gslots_cd;
  long long* d_g1slots = nullptr;
  int2* d_g1sp = nullptr;         // gradient shell pairs (i >= j), slot index order
  int* d_g1pidx = nullptr;        // their indices grouped by class
  int g1_first[NPC] = {}, g1_n[NPC] = {};
  long long* d_gbase = nullptr;
  int* d_gcount = nullptr;
  double* d_gstage = nullptr;
  // per-atom sums: chunks (atom, first slot, slots) and their partial sums
  int4* d_gchunks = nullptr;
  int* d_gcfirst = nullptr;
  double* d_gpart = nullptr;
  int n_gchunk = 0;
  bool tuned = false;
};

namespace {

void free_plan(dft_plan p) {
  if (!p) return;
  for (auto& c : p->combos) {
    (void)hipFree(c.d_q);
    (void)hipFree(c.d_slots);
    (void)hipFree(c.d_segmask);
    (void)hipFree(c.d_lanes);
  }
  void* ptrs[] = {p->d_sh,    p->d_exps,  p->d_coefs,   p->d_pairs,    p->d_shell_dm,  p->d_shell_ao,
                  p->d_rec64, p->d_rec32[0], p->d_ws,      p->d_tg,       p->d_tbase,     p->d_tcount, p->d_tpairs,
                  p->d_atoms, p->d_shpairs, p->d_mol_base, p->d_mol_nbf,   p->d_grec,
                  p->d_gpairs, p->d_g1slots, p->d_gbase, p->d_gcount, p->d_gstage, p->d_recT64, p->d_g1sp, p->d_g1pidx,
                  p->d_gchunks, p->d_gcfirst, p->d_gpart,
                  p->d_pairs32[0], p->d_pairs32[1], p->d_rec32[1], p->d_recT32[0], p->d_recT32[1]};
  for (void* x : ptrs) (void)hipFree(x);
  for (hipEvent_t e : p->prof_ev) (void)hipEventDestroy(e);
  for (hipStream_t x : p->aux) (void)hipStreamDestroy(x);
  for (hipEvent_t e : p->ev_join) (void)hipEventDestroy(e);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  for (auto* x : p->d_gq_swap) (void)hipFree(x);
  for (auto* x : p->d_gq) (void)hipFree(x);
  for (auto* x : p->d_gslots_ab) (void)hipFree(x);
  for (auto* x : p->d_gslots_cd) (void)hipFree(x);
  delete p;
}

long long pair_key(int i, int j) { return ((long long)i << 32) | (unsigned)j; }

// Doubles of a plan's workspace: staging, integral buffers.
long long workspace_len(dft_plan p) { return ws_align(p->stage_len) + p->ibuf_len; }

// The workspace regions of a plan laid out from base.
Workspace workspace_at(dft_plan p, double* base) {
  Workspace w;
  w.stage = base;
  w.ibuf = w.stage + ws_align(p->stage_len);
  return w;
}


RecView rec_view(dft_plan p, int ci, bool f32, int lev = 0);

// Interleaved record blocks per pair class (prec 0: FP64 records and pair
// table, 1 / 2: FP32 levels 0 / 1): element f of primitive pair j of
// class-local pair u at base + (j * rs + f) * n + u, n the class's pair count.
int interleave_records(dft_plan p, int prec) {
  const std::vector<int4>& prs = prec ? p->pairs32[prec - 1] : p->pairs;
  long long* cbase = prec ? p->cls_base32[prec - 1] : p->cls_base;
  std::vector<int4> map;
  long long tot = 0;
  for (int pc = 0; pc < NPC; ++pc) {
    int la, lb;
    pc_moments(pc, la, lb);
    const int rs = rec_size(la, lb), n = p->cls_n[pc], f0 = p->cls_first[pc];
    int mp = 0;
    for (int u = 0; u < n; ++u) mp = std::max(mp, prs[f0 + u].w);
    cbase[pc] = tot;
    for (int u = 0; u < n; ++u) {
      const int4 pr = prs[f0 + u];
      for (int j = 0; j < pr.w; ++j)
        map.push_back(make_int4(pr.z + j * rs, (int)(tot + (long long)j * rs * n + u), rs, n));
    }
    tot += (long long)mp * rs * n;
    if (tot * (long long)sizeof(double) > 0xffffffffLL)
      return fail(DFT_EINVAL, "interleaved pair records exceed 4 GiB");
  }
  int4* dmap = nullptr;
  int rc = upload(&dmap, map);
  if (rc) return rc;
  hipError_t e;
  if (prec == 0) {
    e = hipMalloc(&p->d_recT64, std::max<long long>(tot, 1) * sizeof(double));
    if (e == hipSuccess)
      e = launch_interleave_records(dmap, (int)map.size(), p->d_rec64, nullptr, p->d_recT64, nullptr, nullptr);
  } else {
    float*& t = p->d_recT32[prec - 1];
    e = hipMalloc(&t, std::max<long long>(tot, 1) * sizeof(float));
    if (e == hipSuccess)
      e = launch_interleave_records(dmap, (int)map.size(), nullptr, p->d_rec32[prec - 1], nullptr, t, nullptr);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(dmap);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("record interleave: ") + hipGetErrorString(e));
  return DFT_OK;
}

// Raw contracted blocks for a list of (bra pair, ket pair) in one combo,
// synchronous, host result. By default over the plan's FP64 pair table;
// pairs / rv name another table of the same FP64 records (the primitive
// Schwarz pass) and, for the column-kernel classes, its interleaved view.
int raw_blocks(dft_plan p, int combo, const std::vector<int2>& q, std::vector<double>& out, int blk,
               const int4* pairs = nullptr, const RecView* rvp = nullptr) {
  if (!pairs) pairs = p->d_pairs;
  out.assign(q.size() * (size_t)blk, 0.0);
  if (q.empty()) return DFT_OK;
  int2* dq = nullptr;
  double* dout = nullptr;
  int rc = upload(&dq, q);
  if (rc) return rc;
  hipError_t e = hipMalloc(&dout, out.size() * sizeof(double));
  if (e != hipSuccess) {
    (void)hipFree(dq);
    return fail(DFT_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  if (lane_kernel_combo(combo)) {
    e = launch_eri_raw(combo, dq, (long long)q.size(), pairs, p->d_rec64, dout, p->d_boys, nullptr);
  } else if (generic_combo(combo)) {
    int la, lb, lc, ld;
    pc_moments(combo_pcb(combo), la, lb);
    pc_moments(combo_pck(combo), lc, ld);
    e = launch_eri_generic(la, lb, lc, ld, dq, 0, (int)q.size(), pairs, p->d_rec64, dout, 1, blk, p->d_boys, 1 << 16,
                           false, nullptr);
  } else {
    const RecView rv = rvp ? *rvp : rec_view(p, combo, false);
    e = launch_eri_chunk(combo, false, dq, 0, (int)q.size(), 1, pairs, rv, dout, 1, blk, 0, p->d_boys, 256,
                         1 << 20, false, nullptr);
  }
  if (e == hipSuccess) e = hipMemcpy(out.data(), dout, out.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dq);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("raw ERI blocks: ") + hipGetErrorString(e));
  return DFT_OK;
}

// Input description of a (batch) plan: molecule m owns shells
// [sh_begin[m], sh_begin[m+1]) and atoms [at_begin[m], at_begin[m+1]).
struct Input {
  int nmol;
  std::vector<int> sh_begin, at_begin, nbf;
  const int32_t *L, *NP, *AO;
  const double *XYZ, *EX, *CO, *AXYZ, *AZ;
};

// Host-side phase times of a plan build, printed to stderr when the
// environment variable DFT_PLAN_TIMING is set.
struct PhaseTimer {
  bool on = std::getenv("DFT_PLAN_TIMING") != nullptr;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now(), last = t0;
  std::string out;
  void mark(const char* name) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    char buf[96];
    std::snprintf(buf, sizeof buf, " %s %.2f", name, std::chrono::duration<double, std::milli>(now - last).count());
    out += buf;
    last = now;
  }
  ~PhaseTimer() {
    if (on)
      std::fprintf(stderr, "[dft plan ms]%s total %.2f\n", out.c_str(),
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
};

int build_plan(dft_plan p, const Input& in) {
  PhaseTimer timer;
  const int nsh = p->nsh;
  // ---- molecules: packed matrix bases
  p->nmol = in.nmol;
  p->mol_nbf = in.nbf;
  p->mol_base.assign(in.nmol, 0);
  long long nel = 0;
  for (int m = 0; m < in.nmol; ++m) {
    p->mol_base[m] = nel;
    nel += (long long)in.nbf[m] * in.nbf[m];
    p->max_nbf = std::max(p->max_nbf, in.nbf[m]);
  }
  p->nelem = nel;
  std::vector<int> shell_mol(nsh);
  for (int m = 0; m < in.nmol; ++m)
    for (int i = in.sh_begin[m]; i < in.sh_begin[m + 1]; ++i) shell_mol[i] = m;
  // ---- shells in l-sorted (stable) internal order
  std::vector<int> order(nsh);
  for (int i = 0; i < nsh; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return in.L[a] < in.L[b]; });
  std::vector<long long> prim_in(nsh + 1, 0);
  for (int i = 0; i < nsh; ++i) prim_in[i + 1] = prim_in[i] + in.NP[i];
  p->in2int.assign(nsh, 0);
  p->sh.resize(nsh);
  for (int k = 0; k < nsh; ++k) {
    const int i = order[k];
    const int m = shell_mol[i];
    p->in2int[i] = k;
    ShellDev s;
    s.l = in.L[i];
    s.nprim = in.NP[i];
    s.prim_off = (int)p->exps.size();
    s.ao = in.AO[i];
    s.x = in.XYZ[3 * i];
    s.y = in.XYZ[3 * i + 1];
    s.z = in.XYZ[3 * i + 2];
    s.atom_begin = in.at_begin[m];
    s.atom_end = in.at_begin[m + 1];
    s.nbf = in.nbf[m];
    s.mol = m;
    s.mbase = p->mol_base[m];
    for (int j = 0; j < in.NP[i]; ++j) {
      const double a = in.EX[prim_in[i] + j];
      p->exps.push_back(a);
      p->coefs.push_back(in.CO[prim_in[i] + j] * radial_norm(a, s.l));
    }
    p->sh[k] = s;
  }
  // ---- pairs (within one molecule), sorted by pair class, then molecule,
  // then by surviving primitive-pair count, largest first (stable: (i, j)
  // order among equal counts). The lane kernel runs a wave's bra loop for
  // its lanes' largest count; consecutive bra pairs of similar counts keep
  // those lanes busy (about half the bra iterations idle in (i, j) order).
  auto prim_count = [&](int i, int j) {
    const ShellDev &A = p->sh[i], &B = p->sh[j];
    const double r2 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) + (A.z - B.z) * (A.z - B.z);
    int n = 0;
    for (int ia = 0; ia < A.nprim; ++ia)
      for (int ib = 0; ib < B.nprim; ++ib) {
        const double a = p->exps[A.prim_off + ia], b = p->exps[B.prim_off + ib];
        n += a * b / (a + b) * r2 <= PRIM_CUT;
      }
    return n;
  };
  std::vector<std::vector<int>> mol_shells(in.nmol);
  for (int k = 0; k < nsh; ++k) mol_shells[p->sh[k].mol].push_back(k);
  std::vector<std::vector<int2>> by_pc(NPC);
  for (int m = 0; m < in.nmol; ++m) {
    const auto& ms = mol_shells[m];
    size_t seg[NPC];
    for (int pc = 0; pc < NPC; ++pc) seg[pc] = by_pc[pc].size();
    for (size_t a = 0; a < ms.size(); ++a)
      for (size_t b = 0; b <= a; ++b) {
        const int i = ms[a], j = ms[b];
        by_pc[pc_index(p->sh[i].l, p->sh[j].l)].push_back(make_int2(i, j));
      }
    for (int pc = 0; pc < NPC; ++pc) {
      std::vector<std::pair<int, int2>> v;
      for (size_t k = seg[pc]; k < by_pc[pc].size(); ++k) {
        const int2 ij = by_pc[pc][k];
        v.push_back({prim_count(ij.x, ij.y), ij});
      }
      std::stable_sort(v.begin(), v.end(), [](const std::pair<int, int2>& x, const std::pair<int, int2>& y) {
        return x.first > y.first;
      });
      for (size_t k = 0; k < v.size(); ++k) by_pc[pc][seg[pc] + k] = v[k].second;
    }
  }
  std::vector<int4> pp;       // primitive pairs: sh_a, sh_b, ia, ib
  std::vector<int> pp_off;    // record offsets
  long long roff = 0;
  std::vector<std::vector<std::vector<int>>> class_pairs(in.nmol, std::vector<std::vector<int>>(NPC));
  for (int pc = 0; pc < NPC; ++pc) {
    int la, lb;
    pc_moments(pc, la, lb);
    const int rs = rec_size(la, lb);
    p->cls_first[pc] = (int)p->pairs.size();
    p->cls_n[pc] = (int)by_pc[pc].size();
    for (const int2& ij : by_pc[pc]) {
      const ShellDev &A = p->sh[ij.x], &B = p->sh[ij.y];
      const int id = (int)p->pairs.size();
      const double r2 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) + (A.z - B.z) * (A.z - B.z);
      const int first = (int)p->prim_list.size();
      // primitive-pair screening: exp(-ab/(a+b) R_AB^2) < 1e-17 makes every
      // integral of the primitive pair negligible next to FP64 rounding
      for (int ia = 0; ia < A.nprim; ++ia)
        for (int ib = 0; ib < B.nprim; ++ib) {
          const double a = p->exps[A.prim_off + ia], b = p->exps[B.prim_off + ib];
          if (a * b / (a + b) * r2 > PRIM_CUT) continue;
          p->prim_list.push_back(make_int2(ia, ib));
          pp.push_back(make_int4(ij.x, ij.y, ia, ib));
          pp_off.push_back((int)roff);
          roff += rs;
        }
      p->pairs.push_back(make_int4(ij.x, ij.y, (int)(roff - (long long)rs * ((int)p->prim_list.size() - first)),
                                   (int)p->prim_list.size() - first));
      p->pair_prim_begin.push_back(first);
      p->pair_pc.push_back(pc);
      p->pair_of[pair_key(ij.x, ij.y)] = id;
      class_pairs[A.mol][pc].push_back(id);
      if (roff * (long long)sizeof(double) > 0xffffffffLL) return fail(DFT_EINVAL, "pair records exceed 4 GiB");
    }
  }
  // ---- device data + pair pre-pass
  int rc;
  if ((rc = boys_table_device(&p->d_boys))) return rc;
  if ((rc = upload(&p->d_sh, p->sh))) return rc;
  if ((rc = upload(&p->d_exps, p->exps))) return rc;
  if ((rc = upload(&p->d_coefs, p->coefs))) return rc;
  if ((rc = upload(&p->d_pairs, p->pairs))) return rc;
  std::vector<int4> dm(nsh);
  std::vector<int> ao(nsh);
  for (int k = 0; k < nsh; ++k) {
    const long long mb = p->sh[k].mbase;
    dm[k] = make_int4(p->sh[k].ao, p->sh[k].nbf, (int)(unsigned)(mb & 0xffffffffLL), (int)(mb >> 32));
    ao[k] = p->sh[k].ao;
  }
  if ((rc = upload(&p->d_shell_dm, dm))) return rc;
  if ((rc = upload(&p->d_shell_ao, ao))) return rc;
  if ((rc = upload(&p->d_mol_base, p->mol_base))) return rc;
  if ((rc = upload(&p->d_mol_nbf, p->mol_nbf))) return rc;
  {
    int4* dpp = nullptr;
    int* doff = nullptr;
    if ((rc = upload(&dpp, pp))) return rc;
    if ((rc = upload(&doff, pp_off))) return rc;
    HIP_TRY(hipMalloc(&p->d_rec64, std::max<long long>(roff, 1) * sizeof(double)));
    hipError_t e = launch_pair_records(dpp, (int)pp.size(), doff, p->d_sh, p->d_exps, p->d_coefs, nullptr,
                                       p->d_rec64, nullptr, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(dpp);
    (void)hipFree(doff);
    if (e != hipSuccess) return fail(DFT_EHIP, std::string("pair pre-pass: ") + hipGetErrorString(e));
  }
  if ((rc = interleave_records(p, 0))) return rc;
  timer.mark("pairs+records");
  // ---- atoms and shell pairs for the 1-e kernel
  std::vector<double> at(4 * (size_t)p->natom);
  for (int a = 0; a < p->natom; ++a) {
    at[4 * a] = in.AXYZ[3 * a];
    at[4 * a + 1] = in.AXYZ[3 * a + 1];
    at[4 * a + 2] = in.AXYZ[3 * a + 2];
    at[4 * a + 3] = in.AZ[a];
  }
  if ((rc = upload(&p->d_atoms, at))) return rc;
  std::vector<int2> sp;
  for (int m = 0; m < in.nmol; ++m) {
    const auto& ms = mol_shells[m];
    for (size_t a = 0; a < ms.size(); ++a)
      for (size_t b = 0; b <= a; ++b) sp.push_back(make_int2(ms[a], ms[b]));
  }
  // grouped by class (shells are l-sorted, so the first shell has the larger l)
  {
    std::vector<int2> by;
    for (int pc = 0; pc < NPC; ++pc) {
      p->shp_first[pc] = (int)by.size();
      for (const int2& x : sp) {
        if (p->sh[x.x].l < p->sh[x.y].l) return fail(DFT_EINVAL, "internal: shell pair order");
        if (pc_index(p->sh[x.x].l, p->sh[x.y].l) == pc) by.push_back(x);
      }
      p->shp_n[pc] = (int)by.size() - p->shp_first[pc];
    }
    sp.swap(by);
  }
  p->nshpairs = (int)sp.size();
  if ((rc = upload(&p->d_shpairs, sp))) return rc;
  timer.mark("atoms");
  // ---- Schwarz: diagonal quartets (ab|ab)
  const int npair = (int)p->pairs.size();
  p->Q.assign(npair, 0.0);
  for (int pc = 0; pc < NPC; ++pc) {
    int la, lb;
    pc_moments(pc, la, lb);
    const int nab = ncart_h(la) * ncart_h(lb);
    std::vector<int2> q;
    std::vector<int> ids;
    for (int id = 0; id < npair; ++id)
      if (p->pair_pc[id] == pc) {
        q.push_back(make_int2(id, id));
        ids.push_back(id);
      }
    std::vector<double> blk;
    if ((rc = raw_blocks(p, combo_index(pc, pc), q, blk, nab * nab))) return rc;
    for (size_t k = 0; k < ids.size(); ++k) {
      double mx = 0.0;
      for (int x = 0; x < nab; ++x) mx = std::max(mx, std::fabs(blk[k * nab * nab + (size_t)x * nab + x]));
      p->Q[ids[k]] = std::sqrt(mx);
    }
  }
  timer.mark("schwarz");
  // ---- FP32 primitive lists. q_k = sqrt(max_x (k k|k k)_xx) of each
  // primitive pair k (its records carry c_a c_b) bounds its share of any
  // integral by q_k Q_cd (Cauchy-Schwarz), so dropping the smallest ones of a
  // pair while their sum stays within FP32_PRIM_REL * Q_ab changes every
  // integral of the pair by at most FP32_PRIM_REL times its Schwarz bound.
  // The FP32 records keep the rest (at least one per pair).
  {
    std::vector<int4> ppairs(pp.size());
    for (size_t k = 0; k < pp.size(); ++k) ppairs[k] = make_int4(pp[k].x, pp[k].y, pp_off[k], 1);
    int4* dpp = nullptr;
    if ((rc = upload(&dpp, ppairs))) return rc;
    std::vector<double> qprim(pp.size(), 0.0);
    for (int pc = 0; pc < NPC && rc == DFT_OK; ++pc) {
      const int n = p->cls_n[pc];
      if (n == 0) continue;
      int la, lb;
      pc_moments(pc, la, lb);
      const int nab = ncart_h(la) * ncart_h(lb), rs = rec_size(la, lb), ci = combo_index(pc, pc);
      const int k0 = p->pair_prim_begin[p->cls_first[pc]];
      const int k1 = p->cls_first[pc] + n < npair ? p->pair_prim_begin[p->cls_first[pc] + n] : (int)pp.size();
      // in pieces of at most 32 Ki primitive pairs and 2^24 block elements
      // (each primitive pair's whole diagonal block comes back)
      const int piece = std::max(256, std::min(32768, (1 << 24) / (nab * nab)));
      for (int c0 = k0; c0 < k1 && rc == DFT_OK; c0 += piece) {
        const int nk = std::min(k1 - c0, piece);
        std::vector<int2> q(nk);
        for (int k = 0; k < nk; ++k) q[k] = make_int2(c0 + k, c0 + k);
        std::vector<double> blk;
        if (lane_kernel_combo(ci) || generic_combo(ci)) {
          rc = raw_blocks(p, ci, q, blk, nab * nab, dpp);
        } else {
          // the column kernel reads bra records interleaved over the list
          std::vector<int4> map(nk);
          for (int k = 0; k < nk; ++k) map[k] = make_int4(pp_off[c0 + k], k, rs, nk);
          int4* dmap = nullptr;
          double* tmp = nullptr;
          if ((rc = upload(&dmap, map))) break;
          hipError_t e = hipMalloc(&tmp, (size_t)nk * rs * sizeof(double));
          if (e == hipSuccess) e = launch_interleave_records(dmap, nk, p->d_rec64, nullptr, tmp, nullptr, nullptr);
          if (e == hipSuccess) e = hipDeviceSynchronize();
          if (e == hipSuccess) {
            RecView rv{};
            rv.bra = tmp;
            rv.sb = nk;
            rv.fb = c0;
            rv.ket = p->d_rec64;
            rv.first = 0;
            rv.n = rv.ntot = nk;
            rv.splb = rv.splk = 1;
            rc = raw_blocks(p, ci, q, blk, nab * nab, dpp, &rv);
          } else {
            rc = fail(DFT_EHIP, std::string("primitive Schwarz: ") + hipGetErrorString(e));
          }
          (void)hipFree(dmap);
          (void)hipFree(tmp);
        }
        if (rc) break;
        for (int k = 0; k < nk; ++k) {
          double mx = 0.0;
          for (int x = 0; x < nab; ++x)
            mx = std::max(mx, std::fabs(blk[(size_t)k * nab * nab + (size_t)x * nab + x]));
          qprim[c0 + k] = std::sqrt(mx);
        }
      }
    }
    (void)hipFree(dpp);
    if (rc) return rc;
    // Primitive pairs of each pair in decreasing Schwarz factor, so that the
    // integral loops can stop at the first primitive quartet below the cut;
    // the FP64 records are rewritten in that order with their factors.
    for (int id = 0; id < npair; ++id) {
      const int first = p->pair_prim_begin[id], w = p->pairs[id].w;
      std::vector<int> ord(w);
      for (int k = 0; k < w; ++k) ord[k] = k;
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int u, int v) { return qprim[first + u] > qprim[first + v]; });
      std::vector<int2> pl(w);
      std::vector<int4> pq(w);
      std::vector<double> qq(w);
      for (int k = 0; k < w; ++k) {
        pl[k] = p->prim_list[first + ord[k]];
        pq[k] = pp[first + ord[k]];
        qq[k] = qprim[first + ord[k]];
      }
      for (int k = 0; k < w; ++k) {
        p->prim_list[first + k] = pl[k];
        pp[first + k] = pq[k];
        qprim[first + k] = qq[k];
      }
    }
    {
      int* doff = nullptr;
      double* dq = nullptr;
      dpp = nullptr;
      if (!(rc = upload(&dpp, pp)) && !(rc = upload(&doff, pp_off))) rc = upload(&dq, qprim);
      if (rc) {
        (void)hipFree(dpp);
        (void)hipFree(doff);
        (void)hipFree(dq);
        return rc;
      }
      hipError_t e = launch_pair_records(dpp, (int)pp.size(), doff, p->d_sh, p->d_exps, p->d_coefs, dq,
                                         p->d_rec64, nullptr, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      (void)hipFree(dpp);
      (void)hipFree(doff);
      (void)hipFree(dq);
      if (e != hipSuccess) return fail(DFT_EHIP, std::string("pair pre-pass: ") + hipGetErrorString(e));
      (void)hipFree(p->d_recT64);
      p->d_recT64 = nullptr;
      if ((rc = interleave_records(p, 0))) return rc;
    }
    timer.mark("prim-schwarz+records");
    for (int lev = 0; lev < 2; ++lev) {
      const double eps = lev ? FP32_PRIM_REL / FP32_LEVEL2 : FP32_PRIM_REL;
      std::vector<int4> pp32;
      std::vector<int> pp_off32;
      std::vector<double> q32;
      long long roff32 = 0;
      std::vector<int4>& prs = p->pairs32[lev];
      prs.resize(npair);
      for (int id = 0; id < npair; ++id) {
        const int4 pr = p->pairs[id];
        const int first = p->pair_prim_begin[id];
        int la, lb;
        pc_moments(p->pair_pc[id], la, lb);
        const int rs = rec_size(la, lb);
        std::vector<int> ord(pr.w);
        for (int k = 0; k < pr.w; ++k) ord[k] = k;
        std::sort(ord.begin(), ord.end(), [&](int u, int v) { return qprim[first + u] < qprim[first + v]; });
        std::vector<char> keep(pr.w, 1);
        double dropped = 0.0;
        for (int r = 0; r + 1 < pr.w; ++r) {
          const double qk = qprim[first + ord[r]];
          if (dropped + qk > eps * p->Q[id]) break;
          dropped += qk;
          keep[ord[r]] = 0;
        }
        const long long r0 = roff32;
        int n32 = 0;
        for (int k = 0; k < pr.w; ++k) {
          if (!keep[k]) continue;
          pp32.push_back(pp[first + k]);
          q32.push_back(qprim[first + k]);
          pp_off32.push_back((int)roff32);
          roff32 += rs;
          ++n32;
        }
        prs[id] = make_int4(pr.x, pr.y, (int)r0, n32);
      }
      if ((rc = upload(&p->d_pairs32[lev], prs))) return rc;
      int* doff = nullptr;
      double* dq = nullptr;
      dpp = nullptr;
      if (!(rc = upload(&dpp, pp32)) && !(rc = upload(&doff, pp_off32))) rc = upload(&dq, q32);
      if (rc) {
        (void)hipFree(dpp);
        (void)hipFree(doff);
        (void)hipFree(dq);
        return rc;
      }
      hipError_t e = hipMalloc(&p->d_rec32[lev], std::max<long long>(roff32, 1) * sizeof(float));
      if (e == hipSuccess)
        e = launch_pair_records(dpp, (int)pp32.size(), doff, p->d_sh, p->d_exps, p->d_coefs, dq, nullptr,
                                p->d_rec32[lev], nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      (void)hipFree(dpp);
      (void)hipFree(doff);
      (void)hipFree(dq);
      if (e != hipSuccess) return fail(DFT_EHIP, std::string("FP32 pair pre-pass: ") + hipGetErrorString(e));
      if ((rc = interleave_records(p, 1 + lev))) return rc;
    }
  }
  timer.mark("fp32-records");
  // ---- screened, class-sorted quartet lists (FP64 part, then weak part),
  // molecule-major within each class combination; bra and ket always come
  // from the same molecule
  p->combos.assign(NCOMBO, Combo());
  // lanes a split lane-kernel combination aims at (SPLIT_LANES;
  // DFT_SPLIT_LANES overrides it, for launch-shape studies)
  long long split_lanes = SPLIT_LANES;
  if (const char* e = std::getenv("DFT_SPLIT_LANES")) {
    const long long v = std::atoll(e);
    if (v > 0) split_lanes = v;
  }
  // Lane table of a split lane-kernel combination: ceil(n_b / bpl) entries
  // per quartet (n_b its bra pair's FP64 primitive pairs), all in one wave;
  // entry k of a quartet with ns entries takes bra primitives k, k + ns, ...
  // (the primitives are stored in decreasing Schwarz factor, so every entry
  // gets a like share of those above the primitive-quartet cut);
  // a wave is closed (filled with idle entries that repeat its last
  // quartet, so its ket stays uniform) before a quartet that does not fit,
  // at padding quartets (the end of a molecule's part), where the molecule
  // changes and at the part boundaries n64 and n64 + n32a.
  auto build_lane_table = [&](Combo& c, int bpl) {
    std::vector<int4>& L = c.lanes;
    L.clear();
    auto close = [&]() {
      if (L.empty() || L.size() % WAVE_QUARTETS == 0) return;
      const int4 last = L.back();
      while (L.size() % WAVE_QUARTETS) L.push_back(make_int4(last.x, 0, 0, 1));
    };
    const long long nq = (long long)c.q.size(), b1 = c.n64, b2 = c.n64 + c.n32a;
    int prev_mol = -1;
    for (long long i = 0; i < nq; ++i) {
      if (i == b1) {
        close();
        c.lpos[1] = (long long)L.size();
      }
      if (i == b2) {
        close();
        c.lpos[2] = (long long)L.size();
      }
      const int2 qq = c.q[i];
      if (qq.x < 0) {
        close();
        prev_mol = -1;
        continue;
      }
      const int mol = p->sh[p->pairs[qq.x].x].mol;
      if (mol != prev_mol) close();
      prev_mol = mol;
      const int nl = (std::max(1, p->pairs[qq.x].w) + bpl - 1) / bpl;
      if ((long long)(L.size() % WAVE_QUARTETS) + nl > WAVE_QUARTETS) close();
      for (int k = 0; k < nl; ++k) L.push_back(make_int4((int)i, k, nl, 0));
    }
    close();
    if (b1 >= nq) c.lpos[1] = (long long)L.size();
    if (b2 >= nq) c.lpos[2] = (long long)L.size();
    c.lpos[0] = 0;
    c.lpos[3] = (long long)L.size();
  };
  for (int pcb = 0; pcb < NPC; ++pcb)
    for (int pck = 0; pck <= pcb; ++pck) {
      Combo& c = p->combos[combo_index(pcb, pck)];
      c.pcb = pcb;
      c.pck = pck;
      const bool lane = lane_kernel_combo(combo_index(pcb, pck));
      // the run-time-momentum kernels (f shells) are FP64 only
      const bool gen = generic_combo(combo_index(pcb, pck));
      // per molecule: FP64 part, first and second FP32 level
      std::vector<std::vector<int2>> part[3];
      for (auto& x : part) x.resize(in.nmol);
      // ket-major: consecutive quartets (the lanes of a wave) share the ket
      // pair, whose primitive records the lane kernel reads in its inner loop,
      // so those loads are wave-uniform; the bra pairs of a wave are
      // consecutive pairs of their class (coalesced in the interleaved layout).
      // The (molecule, ket pair) items are screened in contiguous chunks on
      // host threads and the chunks' lists joined in item order.
      std::vector<int2> items;  // (molecule, ket pair)
      long long work = 0;       // bra-ket checks
      for (int m = 0; m < in.nmol; ++m) {
        for (int ik : class_pairs[m][pck]) items.push_back(make_int2(m, ik));
        work += (long long)class_pairs[m][pck].size() * (long long)class_pairs[m][pcb].size();
      }
      // chunks of at least ~2^20 checks (thread start-up costs more below)
      const size_t min_items =
          std::max<size_t>(1, (size_t)((double)items.size() * (double)(1 << 20) / (double)std::max(work, 1LL)));
      struct Chunk {
        std::vector<int2> q[3];
        std::vector<int> mol[3];
        long long nun = 0, prim64 = 0, prim64w = 0, kprod64 = 0, kprod32 = 0;
      };
      std::vector<Chunk> chunks(parallel_chunk_count(items.size(), min_items));
      parallel_chunks(items.size(), min_items, [&](size_t t, size_t i0, size_t i1) {
        Chunk& ch = chunks[t];
        for (size_t it = i0; it < i1; ++it) {
          const int m = items[it].x, ik = items[it].y;
          for (int ib : class_pairs[m][pcb]) {
            if (pcb == pck && ib < ik) continue;
            ++ch.nun;
            const double b = p->Q[ib] * p->Q[ik];
            if (b < p->tau) continue;
            const long long np = (long long)p->pairs[ib].w * p->pairs[ik].w;
            const long long kp = (long long)p->sh[p->pairs[ib].x].nprim * p->sh[p->pairs[ib].y].nprim *
                                 p->sh[p->pairs[ik].x].nprim * p->sh[p->pairs[ik].y].nprim;
            const int k = !gen && b < p->tau32 ? (b < FP32_LEVEL2 * p->tau32 ? 2 : 1) : 0;
            ch.q[k].push_back(make_int2(ib, ik));
            ch.mol[k].push_back(m);
            if (k) {
              ch.prim64w += np;
              ch.kprod32 += kp;
            } else {
              ch.prim64 += np;
              ch.kprod64 += kp;
            }
          }
        }
      });
      for (const Chunk& ch : chunks) {
        p->n_unscreened += ch.nun;
        c.prim64 += ch.prim64;
        c.prim64w += ch.prim64w;
        c.kprod64 += ch.kprod64;
        c.kprod32 += ch.kprod32;
        for (int k = 0; k < 3; ++k)
          for (size_t j = 0; j < ch.q[k].size(); ++j) part[k][ch.mol[k][j]].push_back(ch.q[k][j]);
      }
      long long nreal[3] = {0, 0, 0};
      for (int k = 0; k < 3; ++k)
        for (const auto& v : part[k]) nreal[k] += (long long)v.size();
      // the second FP32 level only where its part is large enough to pay
      // for a launch of its own (otherwise its quartets use the first level's
      // records; the list layout is the same either way)
      const bool level2 = nreal[2] >= FP32_LEVEL2_MIN;
      for (int m = 0; m < in.nmol; ++m) {
        for (const int2& qq : part[1][m]) c.prim32 += (long long)p->pairs32[0][qq.x].w * p->pairs32[0][qq.y].w;
        for (const int2& qq : part[2][m])
          c.prim32 += (long long)p->pairs32[level2 ? 1 : 0][qq.x].w * p->pairs32[level2 ? 1 : 0][qq.y].w;
      }
      c.r64 = nreal[0];
      c.r32a = nreal[1];
      c.r32b = nreal[2];
      // [FP64 parts of all molecules][first FP32 level][second FP32 level];
      // for the lane kernel each molecule's part padded to whole waves with
      // dummies that repeat the last ket pair (so the wave's ket stays uniform)
      for (int k = 0; k < 3; ++k) {
        for (int m = 0; m < in.nmol; ++m) {
          const auto& v = part[k][m];
          c.q.insert(c.q.end(), v.begin(), v.end());
          if (lane && !v.empty())
            while (c.q.size() % WAVE_QUARTETS) c.q.push_back(make_int2(-1, v.back().y));
        }
        if (k == 0) c.n64 = (long long)c.q.size();
        if (k == 1) c.n32a = (long long)c.q.size() - c.n64;
      }
      if (!level2) c.n32a = (long long)c.q.size() - c.n64;
      // primitive split where the class has too few quartets to fill the
      // chip with one lane each: bra chunks first, then ket. Bra chunks go
      // to neighbouring lanes through a lane table (their blocks are summed
      // in the wave); ket chunks stay chunk-major (one slot row each), so the
      // wave keeps one ket primitive range.
      const long long nq = (long long)c.q.size();
      const long long nreal_all = c.r64 + c.r32a + c.r32b;
      long long want = nq > 0 && nq < split_lanes ? (split_lanes + nq - 1) / nq : 1;
      if (nreal_all > 0)  // at least SPLIT_MIN_PRIM primitive quartets per lane
        want = std::min(want, std::max(1LL, (c.kprod64 + c.kprod32) / (nreal_all * SPLIT_MIN_PRIM)));
      if (p->split && lane && want > 1) {
        int nbmax = 0, nkmax = 0;
        for (const int2& qq : c.q) {
          if (qq.x < 0) continue;
          nbmax = std::max(nbmax, p->pairs[qq.x].w);
          nkmax = std::max(nkmax, p->pairs[qq.y].w);
        }
        // bra primitives per lane, at most a wave of lanes per quartet
        const int lpq = (int)std::max<long long>(1, std::min<long long>(want, nbmax));
        int bpl = std::max(1, (nbmax + lpq - 1) / lpq);
        bpl = std::max(bpl, (nbmax + WAVE_QUARTETS - 1) / WAVE_QUARTETS);
        c.splb = std::max(1, (nbmax + bpl - 1) / bpl);
        int splk = (int)std::max<long long>(1, std::min<long long>((want + c.splb - 1) / c.splb, nkmax));
        if (nkmax > 0) {
          const int szk = (nkmax + splk - 1) / splk;  // no empty ket chunk
          splk = (nkmax + szk - 1) / szk;
        }
        c.splk = splk;
        if (c.splb > 1) build_lane_table(c, bpl);
      }
      const long long npos = c.lanes.empty() ? nq * c.splb : (long long)c.lanes.size();
      if (npos * c.splk > 0x7fffffffLL) return fail(DFT_EINVAL, "class combination exceeds 2^31 lanes");
      if (!lane_kernel_combo(combo_index(pcb, pck)) && nq > 0) {
        int la, lb, lc, ld;
        pc_moments(pcb, la, lb);
        pc_moments(pck, lc, ld);
        const long long ni = (long long)ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
        const long long ncd = (long long)ncart_h(lc) * ncart_h(ld);
        // column kernel: split the bra primitives into chunks where the
        // class has too few (quartet, column) lanes to fill the chip (the
        // wave-per-quartet kernel of the f classes is not split)
        if (p->split && !gen && nq * ncd < SPLIT_LANES_HEAVY) {
          int nbmax = 0;
          for (const int2& qq : c.q) nbmax = std::max(nbmax, p->pairs[qq.x].w);
          const long long want = (SPLIT_LANES_HEAVY + nq * ncd - 1) / (nq * ncd);
          c.splb = (int)std::max<long long>(1, std::min<long long>(want, nbmax));
        }
        if (nq * ncd * c.splb > 0x7fffffffLL) return fail(DFT_EINVAL, "class combination exceeds 2^31 lanes");
        c.ibuf_off = p->ibuf_len;
        p->ibuf_len += ws_align(nq * ni * c.splb);
      }
      if ((rc = upload(&c.d_q, c.q))) return rc;
    }
  timer.mark("lists");
  // ---- target-major staging: a target is an ordered shell pair (X, Y) of
  // one molecule in half matrix m (0 = J, 1.. = K per density); its blocks
  // are contiguous and ordered by (combo, quartet), so every sum has a
  // fixed order.
  const int nmat = 1 + p->nk;
  const int nblk = 2 + 4 * p->nk;
  std::vector<long long> mol_toff(in.nmol + 1, 0);  // first target of molecule m (matrix 0)
  std::vector<int> local(nsh);                       // index of the shell within its molecule
  for (int m = 0; m < in.nmol; ++m) {
    const long long ns = (long long)mol_shells[m].size();
    mol_toff[m + 1] = mol_toff[m] + ns * ns;
    for (size_t a = 0; a < mol_shells[m].size(); ++a) local[mol_shells[m][a]] = (int)a;
  }
  const long long tpm = mol_toff[in.nmol];  // targets per matrix kind
  if (nmat * tpm > 0x7fffffffLL) return fail(DFT_EINVAL, "too many reduction targets");
  p->ntarget = (int)(nmat * tpm);
  auto tid = [&](int m, int x, int y) {
    const int mol = p->sh[x].mol;
    const long long ns = (long long)mol_shells[mol].size();
    return (size_t)(m * tpm + mol_toff[mol] + local[x] * ns + local[y]);
  };
  // per pair: its molecule's first target, shells per molecule and the
  // two shells' indices within the molecule (the quartet loops below visit
  // every screened quartet, so the target arithmetic is kept to a few ops)
  struct PairT {
    long long base, ns;
    int x, y;
  };
  std::vector<PairT> pt(p->pairs.size());
  for (size_t id = 0; id < p->pairs.size(); ++id) {
    const int A = p->pairs[id].x, B = p->pairs[id].y, mol = p->sh[A].mol;
    pt[id] = PairT{mol_toff[mol], (long long)mol_shells[mol].size(), local[A], local[B]};
  }
  auto quartet_targets = [&](const int2& qq, size_t* t) {
    const PairT& b = pt[qq.x];
    const PairT& k = pt[qq.y];
    const long long ns = b.ns, base = b.base;
    t[0] = (size_t)(base + b.x * ns + b.y);
    t[1] = (size_t)(base + k.x * ns + k.y);
    for (int m = 0; m < p->nk; ++m) {
      const long long o = (1 + m) * tpm + base;
      t[2 + 4 * m] = (size_t)(o + b.x * ns + k.x);
      t[3 + 4 * m] = (size_t)(o + b.x * ns + k.y);
      t[4 + 4 * m] = (size_t)(o + b.y * ns + k.x);
      t[5 + 4 * m] = (size_t)(o + b.y * ns + k.y);
    }
  };
  std::vector<int> tcount(p->ntarget, 0);
  size_t tq[10];
  // slot rows per quartet: one per primitive chunk for the lane kernel
  // (each chunk writes partial blocks); one for the heavy classes (their
  // digest sums the chunks' integrals first)
  auto slot_rows = [&](size_t ci) {
    const Combo& c = p->combos[ci];
    if (!lane_kernel_combo((int)ci)) return 1;
    return c.lanes.empty() ? c.splb * c.splk : c.splk;
  };
  // Lane-kernel combinations: the lanes of a wave (WAVE_QUARTETS consecutive
  // quartets of one primitive chunk, or WAVE_QUARTETS consecutive entries of
  // the lane table) whose block of kind k goes to the same target form runs;
  // each run is summed across its lanes in the kernel and stored once, by its
  // last lane. The runs depend only on the quartets, so they are found once
  // per wave of the list or table (mask of run starts per block kind, run
  // ends in (kind, lane) order) and shared by every chunk row.
  constexpr size_t NO_TARGET = ~(size_t)0;
  // targets of the 64 positions of wave w (list quartets or lane-table
  // entries; padding and idle positions have none), t[l * nblk + k]
  auto wave_targets = [&](const Combo& c, size_t w, size_t* t) {
    const bool tab = !c.lanes.empty();
    size_t tql[10];
    const size_t w0 = w * WAVE_QUARTETS;
    for (int l = 0; l < WAVE_QUARTETS; ++l) {
      const int4 e = tab ? c.lanes[w0 + l] : make_int4((int)(w0 + l), 0, 0, 0);
      const int2 qq = e.w ? make_int2(-1, 0) : c.q[e.x];
      if (qq.x < 0) {
        for (int k = 0; k < nblk; ++k) t[l * nblk + k] = NO_TARGET;
      } else {
        quartet_targets(qq, tql);
        for (int k = 0; k < nblk; ++k) t[l * nblk + k] = tql[k];
      }
    }
  };
  // Run-start masks of every wave and block kind (host threads over the
  // waves), and the run ends' slot rows counted per target; the slot tables
  // below visit the run ends again in (wave, kind, lane) order. With
  // DFT_PLAN_TIMING the run ends per quartet of each block kind go to stderr.
  std::vector<std::vector<unsigned long long>> masks(p->combos.size());
  for (size_t ci = 0; ci < p->combos.size(); ++ci) {
    const Combo& c = p->combos[ci];
    const int ns = slot_rows(ci);
    if (lane_kernel_combo((int)ci)) {
      const size_t npos = c.lanes.empty() ? c.q.size() : c.lanes.size(), nw = npos / WAVE_QUARTETS;
      std::vector<unsigned long long>& mk = masks[ci];
      mk.assign(nw * nblk, 0);
      std::vector<long long> ends(nblk, 0);
      std::mutex mu;
      const bool one = parallel_chunk_count(nw, 4096) <= 1;  // plain adds on one thread
      parallel_chunks(nw, 4096, [&](size_t, size_t w0, size_t w1) {
        std::vector<size_t> t(WAVE_QUARTETS * nblk);
        std::vector<long long> le(nblk, 0);
        for (size_t w = w0; w < w1; ++w) {
          wave_targets(c, w, t.data());
          for (int k = 0; k < nblk; ++k) {
            unsigned long long m = 0;
            for (int l = 0; l < WAVE_QUARTETS; ++l) {
              const size_t tl = t[l * nblk + k];
              if (l == 0 || tl == NO_TARGET || tl != t[(l - 1) * nblk + k]) m |= 1ull << l;
            }
            mk[w * nblk + k] = m;
            for (int l = 0; l < WAVE_QUARTETS; ++l) {
              const bool last = l == WAVE_QUARTETS - 1 || ((m >> (l + 1)) & 1);
              const size_t tl = t[l * nblk + k];
              if (last && tl != NO_TARGET) {
                if (one) tcount[tl] += ns;
                else __atomic_fetch_add(&tcount[tl], ns, __ATOMIC_RELAXED);
                ++le[k];
              }
            }
          }
        }
        std::lock_guard<std::mutex> g(mu);
        for (int k = 0; k < nblk; ++k) ends[k] += le[k];
      });
      if (timer.on && !c.q.empty()) {
        const long long nreal = c.r64 + c.r32a + c.r32b;
        std::string line = "[dft plan stores/quartet] combo " + std::to_string(ci) + " (" + std::to_string(nreal) +
                           " quartets, " + std::to_string(ns) + " rows):";
        for (int k = 0; k < nblk; ++k) {
          char buf[32];
          std::snprintf(buf, sizeof buf, " %.3f", nreal ? (double)ends[k] / (double)nreal : 0.0);
          line += buf;
        }
        std::fprintf(stderr, "%s\n", line.c_str());
      }
      continue;
    }
    const bool one = parallel_chunk_count(c.q.size(), 1 << 18) <= 1;
    parallel_chunks(c.q.size(), 1 << 18, [&](size_t, size_t i0, size_t i1) {
      size_t tql[10];
      for (size_t i = i0; i < i1; ++i) {
        quartet_targets(c.q[i], tql);
        for (int k = 0; k < nblk; ++k) {
          if (one) tcount[tql[k]] += ns;
          else __atomic_fetch_add(&tcount[tql[k]], ns, __ATOMIC_RELAXED);
        }
      }
    });
  }
  timer.mark("runs");
  std::vector<long long> tbase(p->ntarget, 0), tnext(p->ntarget, 0);
  std::vector<Target> tg(p->ntarget);
  long long cur = 0;
  for (int m = 0; m < nmat; ++m)
    for (int mol = 0; mol < in.nmol; ++mol)
      for (int x : mol_shells[mol])
        for (int y : mol_shells[mol]) {
          const size_t t = tid(m, x, y);
          const int nr = ncart_h(p->sh[x].l), nc = ncart_h(p->sh[y].l);
          const int nbf = in.nbf[mol];
          Target T;
          T.base = m * p->nelem + p->mol_base[mol] + (long long)p->sh[x].ao * nbf + p->sh[y].ao;
          T.nrow = nr;
          T.ncol = nc;
          T.stride = nbf;
          T.pad = m;  // the matrix: 0 = J, 1 + s = K of density s
          tg[t] = T;
          tbase[t] = tnext[t] = cur;
          cur += (long long)tcount[t] * nr * nc;
        }
  p->stage_len = cur;
  // unordered shell pairs {X, Y} of each matrix and molecule: the reduction
  // writes both orientations of a pair's output from its two targets
  {
    std::vector<int2> tp;
    for (int m = 0; m < nmat; ++m)
      for (int mol = 0; mol < in.nmol; ++mol) {
        const std::vector<int>& ms = mol_shells[mol];
        for (size_t i = 0; i < ms.size(); ++i)
          for (size_t j = i; j < ms.size(); ++j)
            tp.push_back(make_int2((int)tid(m, ms[i], ms[j]), (int)tid(m, ms[j], ms[i])));
      }
    p->ntpair = (int)tp.size();
    if ((rc = upload(&p->d_tpairs, tp))) return rc;
  }
  timer.mark("targets");
  // slot tables: block-major for the lane kernel (a wave's loads of one
  // block's slots are contiguous), quartet-major for the heavy-class digest
  // (one wave per quartet)
  for (size_t ci = 0; ci < p->combos.size(); ++ci) {
    Combo& c = p->combos[ci];
    const size_t nq = c.q.size(), ns = (size_t)slot_rows(ci);
    if (lane_kernel_combo((int)ci)) {
      // positions: quartets, or lane-table entries
      const size_t npos = c.lanes.empty() ? nq : c.lanes.size();
      // slot of each run's last lane (-1 for the other lanes), wave by wave
      // in (kind, lane) order; a run owns ns consecutive blocks of its
      // target, one per primitive chunk (the kernel adds ch times the block
      // size). This fixes every target's summation order.
      // Large lists in chunks of waves on host threads: each chunk counts its
      // run ends per target, the counts are turned into each chunk's first
      // slot per target (chunks in order), then each chunk assigns its slots
      // in order: the same slots as one pass over the waves.
      const size_t nw = npos / WAVE_QUARTETS;
      const std::vector<unsigned long long>& mk = masks[ci];
      std::vector<long long> slots(npos * nblk, -1);
      auto visit = [&](size_t w0, size_t w1, auto&& at_end) {
        std::vector<size_t> t(WAVE_QUARTETS * nblk);
        for (size_t w = w0; w < w1; ++w) {
          wave_targets(c, w, t.data());
          for (int k = 0; k < nblk; ++k) {
            const unsigned long long m = mk[w * nblk + k];
            for (int l = 0; l < WAVE_QUARTETS; ++l) {
              const bool last = l == WAVE_QUARTETS - 1 || ((m >> (l + 1)) & 1);
              const size_t tl = t[l * nblk + k];
              if (last && tl != NO_TARGET) at_end((size_t)k * npos + w * WAVE_QUARTETS + l, tl);
            }
          }
        }
      };
      const size_t nchunk = parallel_chunk_count(nw, 4096);
      if (nchunk <= 1) {
        visit(0, nw, [&](size_t pos, size_t tl) {
          slots[pos] = tnext[tl];
          tnext[tl] += (long long)ns * tg[tl].nrow * tg[tl].ncol;
        });
      } else {
        std::vector<std::vector<long long>> cur(nchunk, std::vector<long long>(p->ntarget, 0));
        parallel_chunks(nw, 4096, [&](size_t ch, size_t w0, size_t w1) {
          visit(w0, w1, [&](size_t, size_t tl) { ++cur[ch][tl]; });
        });
        for (size_t tl = 0; tl < (size_t)p->ntarget; ++tl) {
          const long long blk = (long long)ns * tg[tl].nrow * tg[tl].ncol;
          long long off = tnext[tl];
          for (size_t ch = 0; ch < nchunk; ++ch) {
            const long long n = cur[ch][tl];
            cur[ch][tl] = off;
            off += n * blk;
          }
          tnext[tl] = off;
        }
        parallel_chunks(nw, 4096, [&](size_t ch, size_t w0, size_t w1) {
          std::vector<long long>& nx = cur[ch];
          visit(w0, w1, [&](size_t pos, size_t tl) {
            slots[pos] = nx[tl];
            nx[tl] += (long long)ns * tg[tl].nrow * tg[tl].ncol;
          });
        });
      }
      if ((rc = upload(&c.d_slots, slots))) return rc;
      if ((rc = upload(&c.d_segmask, mk))) return rc;
      if (!c.lanes.empty() && (rc = upload(&c.d_lanes, c.lanes))) return rc;
      masks[ci] = std::vector<unsigned long long>();
      continue;
    }
    std::vector<long long> slots(nq * ns * nblk);
    for (size_t ch = 0; ch < ns; ++ch)
      for (size_t i = 0; i < nq; ++i) {
        quartet_targets(c.q[i], tq);
        for (int k = 0; k < nblk; ++k) {
          const Target& T = tg[tq[k]];
          slots[(ch * nq + i) * nblk + k] = tnext[tq[k]];
          tnext[tq[k]] += (long long)T.nrow * T.ncol;
        }
      }
    if ((rc = upload(&c.d_slots, slots))) return rc;
  }
  if ((rc = upload(&p->d_tg, tg))) return rc;
  if ((rc = upload(&p->d_tbase, tbase))) return rc;
  if ((rc = upload(&p->d_tcount, tcount))) return rc;
  timer.mark("slots");
  if (p->own_ws) {
    HIP_TRY(hipMalloc(&p->d_ws, std::max<long long>(workspace_len(p), 1) * sizeof(double)));
    p->ws = workspace_at(p, p->d_ws);
  }
  timer.mark("workspace");
  return DFT_OK;
}

// Largest workgroup of a combination's kernel (lane kernel: per class;
// column kernel: 256, i.e. up to 512 VGPRs per lane).
int block_max(int ci) { return lane_kernel_combo(ci) ? lane_block_max(ci) : 256; }

// Record view of a combination's lane-kernel / column-kernel launch.
// (Also used for the Schwarz diagonals, before the quartet lists exist: the
// bra class comes from the combination index, the list fields only when
// the combination is built.)
RecView rec_view(dft_plan p, int ci, bool f32, int lev) {
  const int pcb = combo_pcb(ci);
  RecView rv;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  const char* t = f32 ? (const char*)p->d_recT32[lev] : (const char*)p->d_recT64;
  rv.bra = t + (f32 ? p->cls_base32[lev][pcb] : p->cls_base[pcb]) * esz;
  rv.sb = p->cls_n[pcb];
  rv.fb = p->cls_first[pcb];
  rv.ket = f32 ? (const void*)p->d_rec32[lev] : (const void*)p->d_rec64;
  rv.first = 0;
  rv.n = 0;
  rv.ntot = 0;
  rv.splb = rv.splk = 1;
  rv.lanes = nullptr;
  if (ci < (int)p->combos.size()) {
    const Combo& c = p->combos[ci];
    rv.ntot = (int)(c.lanes.empty() ? c.q.size() : c.lanes.size());
    rv.splb = c.splb;
    rv.splk = c.splk;
    rv.lanes = c.d_lanes;
  }
  return rv;
}

// One launch over quartets [first, first + n) of one record set (lev -1:
// FP64, 0 / 1: FP32 levels) with the given launch shape. Lane-kernel
// combinations write their J/K blocks directly; heavy combinations write
// integrals to the combination's buffer (launch_digest turns them into J/K
// blocks).
hipError_t launch_level(dft_plan p, const Workspace& ws, size_t ci, int lev, long long first, long long n,
                        const Shape& sh, const double* P, const double* Pa, const double* Pb, hipStream_t s) {
  const Combo& c = p->combos[ci];
  if (n <= 0) return hipSuccess;
  const bool f32 = lev >= 0;
  RecView rv = rec_view(p, (int)ci, f32, f32 ? lev : 0);
  rv.first = (int)first;
  rv.n = (int)n;
  if (rv.lanes) {
    // quartet range -> lane-table range (launches cover whole parts)
    auto pos = [&](long long x) -> long long {
      const long long b[4] = {0, c.n64, c.n64 + c.n32a, (long long)c.q.size()};
      for (int k = 0; k < 4; ++k)
        if (x == b[k]) return c.lpos[k];
      return -1;
    };
    const long long lf = pos(first), le = pos(first + n);
    if (lf < 0 || le < 0) return hipErrorInvalidValue;
    rv.first = (int)lf;
    rv.n = (int)(le - lf);
    if (rv.n <= 0) return hipSuccess;
  }
  const int4* prs = f32 ? p->d_pairs32[lev] : p->d_pairs;  // FP32 launches: that level's primitive lists
  if (lane_kernel_combo((int)ci))
    return launch_eri_jk((int)ci, f32, c.d_q, c.d_slots, c.d_segmask, prs, rv, p->d_shell_dm, P, Pa, Pb, p->nk,
                         ws.stage,
                         p->d_boys, sh.block, sh.cap, s);
  int la, lb, lc, ld;
  pc_moments(c.pcb, la, lb);
  pc_moments(c.pck, lc, ld);
  const long long ni = (long long)ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
  if (generic_combo((int)ci)) {
    if (f32) return hipErrorInvalidValue;  // f classes have no FP32 part
    return launch_eri_generic(la, lb, lc, ld, c.d_q, (int)first, (int)n, p->d_pairs, p->d_rec64,
                              ws.ibuf + c.ibuf_off, 1, ni, p->d_boys, sh.cap * 16, true, s);
  }
  if (!f32 && (p->mfma || c.mfma) && mfma_combo((int)ci))
    return launch_eri_mfma((int)ci, c.d_q, (int)first, (int)n, c.splb, prs, p->d_rec64, ws.ibuf + c.ibuf_off, 1, ni,
                           ni * (long long)c.q.size(), p->d_boys, sh.cap * 4, true, s);
  return launch_eri_chunk((int)ci, f32, c.d_q, (int)first, (int)n, c.splb, prs, rv, ws.ibuf + c.ibuf_off, 1, ni,
                          ni * (long long)c.q.size(), p->d_boys, sh.block, sh.cap, true, s);
}

// The FP64 or FP32 kernel of a class combination on quartets [first, first + n).
hipError_t launch_combo(dft_plan p, const Workspace& ws, size_t ci, bool f32, long long first, long long n,
                        const Shape& sh, const double* P, const double* Pa, const double* Pb, hipStream_t s) {
  const Combo& c = p->combos[ci];
  if (f32) {
    // FP32 quartets [n64, n64 + n32a) use the level-0 records, the rest level 1
    const long long bnd = c.n64 + c.n32a, end = first + n;
    if (first < bnd && end > bnd) {
      hipError_t e = launch_level(p, ws, ci, 0, first, bnd - first, sh, P, Pa, Pb, s);
      return e != hipSuccess ? e : launch_level(p, ws, ci, 1, bnd, end - bnd, sh, P, Pa, Pb, s);
    }
    return launch_level(p, ws, ci, first >= bnd ? 1 : 0, first, n, sh, P, Pa, Pb, s);
  }
  return launch_level(p, ws, ci, -1, first, n, sh, P, Pa, Pb, s);
}

// J/K blocks of a heavy combination from its integral buffer (whole list).
hipError_t launch_digest(dft_plan p, const Workspace& ws, size_t ci, const double* P, const double* Pa,
                         const double* Pb, hipStream_t s) {
  const Combo& c = p->combos[ci];
  if (lane_kernel_combo((int)ci) || c.q.empty()) return hipSuccess;
  int la, lb, lc, ld;
  pc_moments(c.pcb, la, lb);
  pc_moments(c.pck, lc, ld);
  const long long ni = (long long)ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
  if (generic_combo((int)ci))
    return launch_jk_generic(la, lb, lc, ld, c.d_q, c.d_slots, (int)c.q.size(), p->d_pairs, p->d_shell_dm, P, Pa, Pb,
                             p->nk, ws.ibuf + c.ibuf_off, ws.stage, s);
  return launch_jk_digest((int)ci, c.d_q, c.d_slots, (int)c.q.size(), c.splb, ni * (long long)c.q.size(),
                          p->d_pairs, p->d_shell_dm, P, Pa, Pb, p->nk, ws.ibuf + c.ibuf_off, ws.stage, s);
}

// The whole Fock build: class launches, reduction, symmetrisation. The class
// launches write disjoint staging slots, so they run concurrently: each goes
// to the least-loaded of the caller's stream and the plan's auxiliary
// streams (greedy by estimated cost, largest first), which fork from and
// join back into the caller's stream by events (also under graph capture).
int jk_enqueue(dft_plan p, const Workspace& ws, const double* P, const double* Pa, const double* Pb, double* J,
               double* Ka, double* Kb, int mode, hipStream_t s, bool profile) {
  const int nslot = 2 * (int)p->combos.size() + 2;
  hipEvent_t* ev = nullptr;
  if (profile && p->prof_used < p->prof_steps) ev = &p->prof_ev[(size_t)p->prof_used++ * nslot * 2];
  auto mark = [&](int slot, int end, hipStream_t st) {
    if (ev) (void)hipEventRecord(ev[2 * slot + end], st);
  };
  struct Launch {
    int ci;
    bool f32;
    long long first, n;
    double cost;
  };
  std::vector<Launch> ls;
  for (size_t ci = 0; ci < p->combos.size(); ++ci) {
    const Combo& c = p->combos[ci];
    const long long ntot = (long long)c.q.size();
    if (ntot == 0) continue;
    // FP64 mode: one launch over the whole list; mixed: FP64 part, then the
    // weak part through the FP32 kernel (classes whose FP32 launch does not
    // pay run as in FP64 mode)
    const bool f32 = mode == DFT_MIXED && c.use32;
    const long long n1 = f32 ? c.n64 : ntot;
    int la, lb, lc, ld;
    pc_moments(c.pcb, la, lb);
    pc_moments(c.pck, lc, ld);
    const double w = 16.0 + ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);  // per primitive quartet
    if (!lane_kernel_combo((int)ci)) {
      // heavy combination: integral launches and digest in order on one stream
      ls.push_back({(int)ci, false, 0, ntot, w * (c.prim64 + (f32 ? c.prim32 : c.prim64w))});
      continue;
    }
    if (n1 > 0) ls.push_back({(int)ci, false, 0, n1, w * (f32 ? c.prim64 : c.prim64 + c.prim64w)});
    if (f32 && ntot > c.n64) ls.push_back({(int)ci, true, c.n64, ntot - c.n64, 0.7 * w * c.prim32});
  }
  std::stable_sort(ls.begin(), ls.end(), [](const Launch& a, const Launch& b) { return a.cost > b.cost; });
  // slots without a launch read as zero
  for (int slot = 0; slot < nslot - 2; ++slot) {
    mark(slot, 0, s);
    mark(slot, 1, s);
  }
  std::vector<hipStream_t> st(1, s);
  if (p->concurrent)
    for (hipStream_t a : p->aux) st.push_back(a);
  const bool fork = ls.size() > 1 && st.size() > 1;
  if (fork) {
    HIP_TRY(hipEventRecord(p->ev_fork, s));
    for (size_t k = 1; k < st.size(); ++k) HIP_TRY(hipStreamWaitEvent(st[k], p->ev_fork, 0));
  }
  std::vector<double> load(st.size(), 0.0);
  for (const Launch& L : ls) {
    size_t k = 0;
    if (fork)
      for (size_t x = 1; x < st.size(); ++x)
        if (load[x] < load[k]) k = x;
    load[k] += L.cost;
    const Combo& c = p->combos[L.ci];
    if (!lane_kernel_combo(L.ci)) {
      const bool f32 = mode == DFT_MIXED && c.use32;
      const long long n1 = f32 ? c.n64 : (long long)c.q.size();
      mark(2 * L.ci, 0, st[k]);
      hipError_t e = launch_combo(p, ws, L.ci, false, 0, n1, c.shape64, P, Pa, Pb, st[k]);
      mark(2 * L.ci, 1, st[k]);
      mark(2 * L.ci + 1, 0, st[k]);
      if (e == hipSuccess && f32)
        e = launch_combo(p, ws, L.ci, true, c.n64, (long long)c.q.size() - c.n64, c.shape32, P, Pa, Pb, st[k]);
      if (e == hipSuccess) e = launch_digest(p, ws, L.ci, P, Pa, Pb, st[k]);
      mark(2 * L.ci + 1, 1, st[k]);
      if (e != hipSuccess) return fail(DFT_EHIP, std::string("eri column / digest: ") + hipGetErrorString(e));
      continue;
    }
    const int slot = 2 * L.ci + (L.f32 ? 1 : 0);
    mark(slot, 0, st[k]);
    hipError_t e =
        launch_combo(p, ws, L.ci, L.f32, L.first, L.n, L.f32 ? c.shape32 : c.shape64, P, Pa, Pb, st[k]);
    if (e != hipSuccess)
      return fail(DFT_EHIP, std::string(L.f32 ? "eri_jk fp32: " : "eri_jk fp64: ") + hipGetErrorString(e));
    mark(slot, 1, st[k]);
  }
  if (fork)
    for (size_t k = 1; k < st.size(); ++k) {
      HIP_TRY(hipEventRecord(p->ev_join[k - 1], st[k]));
      HIP_TRY(hipStreamWaitEvent(s, p->ev_join[k - 1], 0));
    }
  mark(nslot - 2, 0, s);
  // the reduction writes J = Jh + Jh^T and K likewise directly (the last
  // profile slot, once the separate symmetrisation, stays empty)
  hipError_t e = launch_reduce_sym(p->ntpair, p->d_tpairs, p->d_tg, p->d_tbase, p->d_tcount, ws.stage, p->stage_len,
                                   p->nelem, J, Ka, Kb, s);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("reduce: ") + hipGetErrorString(e));
  mark(nslot - 2, 1, s);
  mark(nslot - 1, 0, s);
  mark(nslot - 1, 1, s);
  return DFT_OK;
}

int check_jk_args(dft_plan p, const double* P, const double* Pa, const double* Pb, double* J, double* Ka, double* Kb,
                  int mode) {
  if (!p) return fail(DFT_EINVAL, "plan is NULL");
  if (!P || !Pa || !J || !Ka) return fail(DFT_EINVAL, "P, Pa, J, Ka must be non-NULL");
  if ((p->nk == 2) != (Pb != nullptr) || (p->nk == 2) != (Kb != nullptr))
    return fail(DFT_EINVAL, "Pb/Kb must be given exactly for a UHF plan");
  if (mode != DFT_FP64 && mode != DFT_MIXED) return fail(DFT_EINVAL, "mode must be DFT_FP64 or DFT_MIXED");
  return DFT_OK;
}

int need_own_workspace(dft_plan p) {
  if (!p->own_ws)
    return fail(DFT_EINVAL, "plan was built with DFT_PLAN_STATIC_WORKSPACE: it has no workspace of its own, "
                            "build its J/K with dft_jk_static");
  return DFT_OK;
}

// Gradient set-up: extended pair records, swapped quartet lists and the
// per-atom slot map (2-e: four centres per quartet; 1-e: three per
// (shell pair, nucleus) block). Slots of one atom are contiguous and ordered
// (2-e AB, 2-e CD per combination, then 1-e), so the sum order is fixed.
int grad_prepare(dft_plan p, const int32_t* shell_atom) {
  PhaseTimer timer;
  const int nsh = p->nsh;
  p->grad_atom.assign(nsh, 0);
  for (int i = 0; i < nsh; ++i) {
    const int a = shell_atom[i];
    if (a < 0 || a >= p->natom) return fail(DFT_EINVAL, "shell_atom out of range");
    p->grad_atom[p->in2int[i]] = a;
  }
  for (const auto& s : p->sh)
    if (s.l > GRAD_LMAX) return fail(DFT_EUNSUPPORTED, "analytic gradients are built for s and p shells only");
  std::vector<int4> pp, gp(p->pairs.size());
  std::vector<int> off;
  long long roff = 0;
  for (size_t id = 0; id < p->pairs.size(); ++id) {
    const int4 pr = p->pairs[id];
    const ShellDev &A = p->sh[pr.x], &B = p->sh[pr.y];
    gp[id] = make_int4(pr.x, pr.y, (int)roff, pr.w);
    const int rs = 6 + 3 * ne_axis(A.l + 1, B.l + 1);
    // the same screened primitive pairs as the J/K records
    for (int k = 0; k < pr.w; ++k) {
      const int2 ab = p->prim_list[p->pair_prim_begin[id] + k];
      pp.push_back(make_int4(pr.x, pr.y, ab.x, ab.y));
      off.push_back((int)roff);
      roff += rs;
    }
  }
  int rc;
  if ((rc = upload(&p->d_gpairs, gp))) return rc;
  HIP_TRY(hipMalloc(&p->d_grec, std::max<long long>(roff, 1) * sizeof(double)));
  {
    int4* dpp = nullptr;
    int* doff = nullptr;
    if ((rc = upload(&dpp, pp))) return rc;
    if ((rc = upload(&doff, off))) return rc;
    hipError_t e = launch_grad_records(dpp, (int)pp.size(), doff, p->d_sh, p->d_exps, p->d_coefs, p->d_grec, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(dpp);
    (void)hipFree(doff);
    if (e != hipSuccess) return fail(DFT_EHIP, std::string("gradient records: ") + hipGetErrorString(e));
  }
  timer.mark("grad-records");
  const int natom = p->natom;
  std::vector<int> cnt(natom, 0);
  auto atom_of = [&](int pid, int which) { return p->grad_atom[which == 0 ? p->pairs[pid].x : p->pairs[pid].y]; };
  for (const Combo& c : p->combos)
    for (const int2& q : c.q) {
      if (q.x < 0) continue;  // padding of the lane-kernel lists
      ++cnt[atom_of(q.x, 0)];
      ++cnt[atom_of(q.x, 1)];
      ++cnt[atom_of(q.y, 0)];
      ++cnt[atom_of(q.y, 1)];
    }
  const int nsp = p->nshpairs;
  std::vector<int2> sp;
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j <= i; ++j) sp.push_back(make_int2(i, j));
  for (int k = 0; k < nsp; ++k)
    for (int cn = 0; cn < natom; ++cn) {
      ++cnt[p->grad_atom[sp[k].x]];
      ++cnt[p->grad_atom[sp[k].y]];
      ++cnt[cn];
    }
  std::vector<long long> base(natom), next(natom);
  long long cur = 0;
  for (int a = 0; a < natom; ++a) {
    base[a] = next[a] = cur;
    cur += 3LL * cnt[a];
  }
  auto slot = [&](int atom) {
    const long long s = next[atom];
    next[atom] += 3;
    return s;
  };
  for (Combo& c : p->combos) {
    // the combination's quartets without the padding
    std::vector<int2> cq;
    for (const int2& q : c.q)
      if (q.x >= 0) cq.push_back(q);
    std::vector<long long> sab(2 * cq.size()), scd(2 * cq.size());
    std::vector<int2> sw(cq.size());
    for (size_t i = 0; i < cq.size(); ++i) {
      sab[2 * i] = slot(atom_of(cq[i].x, 0));
      sab[2 * i + 1] = slot(atom_of(cq[i].x, 1));
    }
    for (size_t i = 0; i < cq.size(); ++i) {
      scd[2 * i] = slot(atom_of(cq[i].y, 0));
      scd[2 * i + 1] = slot(atom_of(cq[i].y, 1));
      sw[i] = make_int2(cq[i].y, cq[i].x);
    }
    int2 *dfw = nullptr, *dsw = nullptr;
    long long *dab = nullptr, *dcd = nullptr;
    if ((rc = upload(&dfw, cq))) return rc;
    if ((rc = upload(&dsw, sw))) return rc;
    if ((rc = upload(&dab, sab))) return rc;
    if ((rc = upload(&dcd, scd))) return rc;
    p->d_gq.push_back(dfw);
    p->gq_n.push_back((long long)cq.size());
    p->d_gq_swap.push_back(dsw);
    p->d_gslots_ab.push_back(dab);
    p->d_gslots_cd.push_back(dcd);
  }
  timer.mark("grad-2e-slots");
  std::vector<long long> s1(3 * (size_t)nsp * natom);
  for (int k = 0; k < nsp; ++k)
    for (int cn = 0; cn < natom; ++cn) {
      const size_t b = 3 * ((size_t)k * natom + cn);
      s1[b] = slot(p->grad_atom[sp[k].x]);
      s1[b + 1] = slot(p->grad_atom[sp[k].y]);
      s1[b + 2] = slot(cn);
    }
  if ((rc = upload(&p->d_g1slots, s1))) return rc;
  // 1-e derivative kernels: shell pairs grouped by class
  {
    std::vector<int> pidx;
    for (int pc = 0; pc < NPC; ++pc) {
      p->g1_first[pc] = (int)pidx.size();
      for (int k = 0; k < nsp; ++k)
        if (pc_index(p->sh[sp[k].x].l, p->sh[sp[k].y].l) == pc) pidx.push_back(k);
      p->g1_n[pc] = (int)pidx.size() - p->g1_first[pc];
    }
    if ((rc = upload(&p->d_g1sp, sp))) return rc;
    if ((rc = upload(&p->d_g1pidx, pidx))) return rc;
  }
  if ((rc = upload(&p->d_gbase, base))) return rc;
  if ((rc = upload(&p->d_gcount, cnt))) return rc;
  HIP_TRY(hipMalloc(&p->d_gstage, std::max<long long>(cur, 1) * sizeof(double)));
  {
    std::vector<int4> chunks;
    std::vector<int> cfirst(natom + 1, 0);
    for (int a = 0; a < natom; ++a) {
      cfirst[a] = (int)chunks.size();
      for (int k = 0; k < cnt[a]; k += GRAD_CHUNK) chunks.push_back(make_int4(a, k, std::min(GRAD_CHUNK, cnt[a] - k), 0));
    }
    cfirst[natom] = (int)chunks.size();
    p->n_gchunk = (int)chunks.size();
    if ((rc = upload(&p->d_gchunks, chunks))) return rc;
    if ((rc = upload(&p->d_gcfirst, cfirst))) return rc;
    HIP_TRY(hipMalloc(&p->d_gpart, std::max<long long>(3LL * p->n_gchunk, 1) * sizeof(double)));
  }
  timer.mark("grad-1e-slots");
  p->grad_ready = true;
  return DFT_OK;
}

// auxiliary streams and fork/join events on the current device
int create_streams(dft_plan p) {
  HIP_TRY(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
  for (int k = 0; k < JK_AUX; ++k) {
    hipStream_t x;
    HIP_TRY(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    p->aux.push_back(x);
    hipEvent_t e;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    p->ev_join.push_back(e);
  }
  return DFT_OK;
}

int validate_and_build(int nmol, const int32_t* mol_nsh, const int32_t* mol_natom, const int32_t* mol_nbf,
                       const int32_t* shell_l, const int32_t* shell_nprim, const int32_t* shell_ao,
                       const double* shell_xyz, const double* prim_exp, const double* prim_coef, const double* atom_xyz,
                       const double* atom_charge, double schwarz_tau, double fp32_tau, int flags, dft_plan* out) {
  if (!out) return fail(DFT_EINVAL, "out is NULL");
  *out = nullptr;
  if (nmol <= 0 || !mol_nsh || !mol_natom || !mol_nbf) return fail(DFT_EINVAL, "bad molecule arrays");
  if (!shell_l || !shell_nprim || !shell_ao || !shell_xyz || !prim_exp || !prim_coef)
    return fail(DFT_EINVAL, "NULL shell array");
  if (!(schwarz_tau >= 0.0) || !(fp32_tau >= 0.0)) return fail(DFT_EINVAL, "thresholds must be >= 0");
  Input in;
  in.nmol = nmol;
  in.sh_begin.assign(nmol + 1, 0);
  in.at_begin.assign(nmol + 1, 0);
  in.nbf.assign(nmol, 0);
  for (int m = 0; m < nmol; ++m) {
    if (mol_nsh[m] <= 0 || mol_nbf[m] <= 0 || mol_natom[m] < 0)
      return fail(DFT_EINVAL, "molecule " + std::to_string(m) + ": nsh, nbf must be positive, natom >= 0");
    in.sh_begin[m + 1] = in.sh_begin[m] + mol_nsh[m];
    in.at_begin[m + 1] = in.at_begin[m] + mol_natom[m];
    in.nbf[m] = mol_nbf[m];
  }
  const int nsh = in.sh_begin[nmol], natom = in.at_begin[nmol];
  if (natom > 0 && (!atom_xyz || !atom_charge)) return fail(DFT_EINVAL, "NULL atom array");
  for (int m = 0; m < nmol; ++m)
    for (int i = in.sh_begin[m]; i < in.sh_begin[m + 1]; ++i) {
      if (shell_l[i] < 0 || shell_l[i] > LMAX)
        return fail(DFT_EUNSUPPORTED, "shell " + std::to_string(i) + ": angular momentum " +
                                          std::to_string(shell_l[i]) + " not built (max " + std::to_string(LMAX) +
                                          ")");
      if (shell_nprim[i] <= 0) return fail(DFT_EINVAL, "shell " + std::to_string(i) + ": nprim <= 0");
      if (shell_ao[i] < 0 || shell_ao[i] + ncart_h(shell_l[i]) > mol_nbf[m])
        return fail(DFT_EINVAL, "shell " + std::to_string(i) + ": AO range outside [0, nbf)");
    }
  in.L = shell_l;
  in.NP = shell_nprim;
  in.AO = shell_ao;
  in.XYZ = shell_xyz;
  in.EX = prim_exp;
  in.CO = prim_coef;
  in.AXYZ = atom_xyz;
  in.AZ = atom_charge;
  dft_plan p = new dft_plan_s();
  p->nsh = nsh;
  p->natom = natom;
  p->nk = (flags & DFT_PLAN_UHF) ? 2 : 1;
  p->split = !(flags & DFT_PLAN_NOSPLIT);
  p->own_ws = !(flags & DFT_PLAN_STATIC_WORKSPACE);
  p->mfma = (flags & DFT_PLAN_MFMA) != 0;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(DFT_EHIP, "hipGetDevice failed");
  }
  p->tau = schwarz_tau;
  p->tau32 = fp32_tau;
  int rc = build_plan(p, in);
  if (!rc && !(flags & DFT_PLAN_SERIAL)) rc = create_streams(p);
  if (!rc) {
    long long nq = 0;
    for (const Combo& c : p->combos) nq += (long long)c.q.size();
    p->concurrent = !(flags & DFT_PLAN_SERIAL) && nq < CONCURRENT_QUARTETS;
  }
  if (rc) {
    free_plan(p);
    return rc;
  }
  *out = p;
  return DFT_OK;
}

}  // namespace

// Static device scratch of dft_jk_static: one copy per GPU, held until the
// process ends. Host code never uses this name as an address (on the host it
// is only a handle of the symbol): each device's copy is reached through the
// pointer hipGetSymbolAddress returns for it (scratch_on_current_device).
static __device__ __align__(256) double scratch[DFT_STATIC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr long long SCRATCH_DOUBLES = DFT_STATIC_SCRATCH_BYTES / (long long)sizeof(double);
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[SCRATCH_MAX_DEVICES] = {};

// Device address of the current device's copy of the scratch, looked up once
// per device (with that device current) and kept.
int scratch_on_current_device(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(DFT_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(DFT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch[dev];
  return DFT_OK;
}

}  // namespace

extern "C" {

int dft_version(void) { return 200; }

const char* dft_last_error(void) { return g_err.c_str(); }

int dft_plan_create(int nsh, const int32_t* shell_l, const int32_t* shell_nprim, const int32_t* shell_ao,
                    const double* shell_xyz, const double* prim_exp, const double* prim_coef, int natom,
                    const double* atom_xyz, const double* atom_charge, int nbf, double schwarz_tau,
                    double fp32_tau, int flags, dft_plan* out) {
  if (nsh <= 0 || nbf <= 0 || natom < 0) {
    if (out) *out = nullptr;
    return fail(DFT_EINVAL, "nsh, nbf must be positive, natom >= 0");
  }
  const int32_t ns = nsh, na = natom, nb = nbf;
  return validate_and_build(1, &ns, &na, &nb, shell_l, shell_nprim, shell_ao, shell_xyz, prim_exp, prim_coef,
                            atom_xyz, atom_charge, schwarz_tau, fp32_tau, flags, out);
}

int dft_plan_create_batch(int nmol, const int32_t* mol_nsh, const int32_t* mol_natom, const int32_t* mol_nbf,
                          const int32_t* shell_l, const int32_t* shell_nprim, const int32_t* shell_ao,
                          const double* shell_xyz, const double* prim_exp, const double* prim_coef,
                          const double* atom_xyz, const double* atom_charge, double schwarz_tau, double fp32_tau,
                          int flags, dft_plan* out) {
  return validate_and_build(nmol, mol_nsh, mol_natom, mol_nbf, shell_l, shell_nprim, shell_ao, shell_xyz, prim_exp,
                            prim_coef, atom_xyz, atom_charge, schwarz_tau, fp32_tau, flags, out);
}

int dft_plan_destroy(dft_plan plan) {
  free_plan(plan);
  return DFT_OK;
}

int dft_plan_stats(dft_plan p, int64_t* counts, int ncombo, int* ncombo_out, int64_t* n_unscreened,
                   int64_t* staging_bytes) {
  if (!p) return fail(DFT_EINVAL, "plan is NULL");
  if (ncombo_out) *ncombo_out = (int)p->combos.size();
  if (counts)
    for (int c = 0; c < std::min<int>(ncombo, (int)p->combos.size()); ++c) {
      const Combo& cb = p->combos[c];
      counts[DFT_STATS_COLS * c] = lane_kernel_combo(c) ? cb.r64 : (int64_t)cb.n64;
      counts[DFT_STATS_COLS * c + 1] =
          lane_kernel_combo(c) ? cb.r32a + cb.r32b : (int64_t)cb.q.size() - cb.n64;
      counts[DFT_STATS_COLS * c + 2] = cb.prim64;
      counts[DFT_STATS_COLS * c + 3] = cb.prim32;
      counts[DFT_STATS_COLS * c + 4] = cb.kprod64;
      counts[DFT_STATS_COLS * c + 5] = cb.kprod32;
      counts[DFT_STATS_COLS * c + 6] =
          lane_kernel_combo(c) ? (cb.n64 + cb.n32a < (long long)cb.q.size() ? cb.r32b : 0)
                               : (int64_t)cb.q.size() - cb.n64 - cb.n32a;
      counts[DFT_STATS_COLS * c + 7] = cb.prim64w;
    }
  if (n_unscreened) *n_unscreened = p->n_unscreened;
  if (staging_bytes) *staging_bytes = workspace_len(p) * (int64_t)sizeof(double);
  return DFT_OK;
}

int dft_schwarz(dft_plan p, double* Q) {
  if (!p || !Q) return fail(DFT_EINVAL, "NULL argument");
  if (p->nmol != 1) return fail(DFT_EINVAL, "dft_schwarz takes a single-molecule plan");
  const int n = p->nsh;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      int a = p->in2int[i], b = p->in2int[j];
      if (a < b) std::swap(a, b);
      Q[(size_t)i * n + j] = p->Q[p->pair_of.at(pair_key(a, b))];
    }
  return DFT_OK;
}

int dft_jk(dft_plan p, const double* P, const double* Pa, const double* Pb, double* J, double* Ka, double* Kb,
           int mode, void* stream) {
  int rc = check_jk_args(p, P, Pa, Pb, J, Ka, Kb, mode);
  if (!rc) rc = need_own_workspace(p);
  if (rc) return rc;
  return jk_enqueue(p, p->ws, P, Pa, Pb, J, Ka, Kb, mode, (hipStream_t)stream, true);
}

int dft_jk_static(dft_plan p, const double* P, const double* Pa, const double* Pb, double* J, double* Ka, double* Kb,
                  int mode, void* stream) {
  int rc = check_jk_args(p, P, Pa, Pb, J, Ka, Kb, mode);
  if (rc) return rc;
  const long long need = workspace_len(p);
  if (need > SCRATCH_DOUBLES)
    return fail(DFT_EINVAL, "dft_jk_static: the plan needs " + std::to_string(need * (long long)sizeof(double)) +
                                " bytes of workspace, the static scratch holds " +
                                std::to_string(DFT_STATIC_SCRATCH_BYTES) + " (use dft_jk)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != p->device) return fail(DFT_EINVAL, "dft_jk_static: the plan's device must be the current device");
  double* base = nullptr;
  if ((rc = scratch_on_current_device(&base))) return rc;
  return jk_enqueue(p, workspace_at(p, base), P, Pa, Pb, J, Ka, Kb, mode, (hipStream_t)stream, true);
}

int dft_jk_graph_create(dft_plan p, const double* P, const double* Pa, const double* Pb, double* J, double* Ka,
                        double* Kb, int mode, dft_graph* out) {
  if (!out) return fail(DFT_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_jk_args(p, P, Pa, Pb, J, Ka, Kb, mode);
  if (!rc) rc = need_own_workspace(p);
  if (rc) return rc;
  hipStream_t cs = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = jk_enqueue(p, p->ws, P, Pa, Pb, J, Ka, Kb, mode, cs, false);
    hipError_t e2 = hipStreamEndCapture(cs, &graph);
    if (rc == DFT_OK && e2 != hipSuccess) e = e2;
  }
  hipGraphExec_t exec = nullptr;
  if (e == hipSuccess && rc == DFT_OK) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(cs);
  if (rc) return rc;
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  *out = (dft_graph)exec;
  return DFT_OK;
}

int dft_graph_launch(dft_graph g, void* stream) {
  if (!g) return fail(DFT_EINVAL, "graph is NULL");
  HIP_TRY(hipGraphLaunch((hipGraphExec_t)g, (hipStream_t)stream));
  return DFT_OK;
}

int dft_graph_destroy(dft_graph g) {
  if (g) (void)hipGraphExecDestroy((hipGraphExec_t)g);
  return DFT_OK;
}

int dft_plan_tune(dft_plan p, const double* P, const double* Pa, const double* Pb, int reps, int combo, int* report,
                  int report_cap, void* stream) {
  if (!p || !P || !Pa) return fail(DFT_EINVAL, "NULL argument");
  if ((p->nk == 2) != (Pb != nullptr)) return fail(DFT_EINVAL, "Pb must be given exactly for a UHF plan");
  int rc0 = need_own_workspace(p);
  if (rc0) return rc0;
  if (reps < 1) reps = 1;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  // candidate launch shapes: workgroup size x grid cap (lane and column kernels)
  std::vector<Shape> lane_c;
  for (int b : {64, 128, 256, 512})
    for (int cap : {256 * 2, 256 * 4, 256 * 8, 256 * 16, 256 * 32}) {
      Shape sh;
      sh.wave = 0;
      sh.block = b;
      sh.cap = cap;
      lane_c.push_back(sh);
    }
  int rc = DFT_OK;
  for (size_t ci = 0; ci < p->combos.size() && rc == DFT_OK; ++ci) {
    if (combo == -2 || (combo >= 0 && (int)ci != combo)) continue;
    if (generic_combo((int)ci)) continue;  // one wave per quartet: no launch shape to choose
    Combo& c = p->combos[ci];
    const long long ntot = (long long)c.q.size();
    float t_all = 0.f, t_32 = 0.f;
    for (int part = 0; part < 2 && rc == DFT_OK; ++part) {
      // part 0: whole list in FP64 (used by the FP64 mode and, for its FP64
      // part, by the mixed mode); part 1: weak part in FP32
      const long long first = part ? c.n64 : 0, n = part ? ntot - c.n64 : ntot;
      if (n <= 0) continue;
      std::vector<Shape> cand;
      for (const Shape& sh : lane_c)
        if (sh.block <= block_max((int)ci)) cand.push_back(sh);
      float best = 1e30f;
      Shape best_sh = part ? c.shape32 : c.shape64;
      if (part == 0) c.mfma = false;  // shapes are searched on the column kernel
      for (const Shape& sh : cand) {
        hipError_t e = launch_combo(p, p->ws, ci, part == 1, first, n, sh, P, Pa, Pb, s);  // warm-up
        std::vector<float> t;
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
          // a shape already 2x slower than the best one is not timed again
          if (r > 0 && !t.empty() && t[0] > 2.0f * best) break;
          (void)hipEventRecord(e0, s);
          e = launch_combo(p, p->ws, ci, part == 1, first, n, sh, P, Pa, Pb, s);
          (void)hipEventRecord(e1, s);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          t.push_back(ms);
        }
        if (e != hipSuccess) {
          rc = fail(DFT_EHIP, std::string("tune: ") + hipGetErrorString(e));
          break;
        }
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (med < best) {
          best = med;
          best_sh = sh;
        }
      }
      if (part == 0 && rc == DFT_OK && !p->mfma && mfma_combo((int)ci)) {
        // the matrix-core variant of the FP64 part, kept where it is faster
        c.mfma = true;
        hipError_t e = launch_combo(p, p->ws, ci, false, first, n, best_sh, P, Pa, Pb, s);
        std::vector<float> t;
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
          (void)hipEventRecord(e0, s);
          e = launch_combo(p, p->ws, ci, false, first, n, best_sh, P, Pa, Pb, s);
          (void)hipEventRecord(e1, s);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          t.push_back(ms);
        }
        if (e != hipSuccess) {
          rc = fail(DFT_EHIP, std::string("tune (mfma): ") + hipGetErrorString(e));
        } else {
          std::sort(t.begin(), t.end());
          c.mfma = t[t.size() / 2] < best;
          if (c.mfma) best = t[t.size() / 2];
        }
      }
      if (part == 0) {
        c.shape64 = best_sh;
        c.shape32 = best_sh;  // default for the FP32 launch until tuned below
        t_all = best;
      } else {
        c.shape32 = best_sh;
        t_32 = best;
      }
    }
    // mixed mode: FP64 part + FP32 weak part against one FP64 launch over
    // the whole list (the FP32 kernel runs FP32 FMAs at the FP64 rate, so
    // for some classes the split does not pay)
    if (rc == DFT_OK && ntot > c.n64) {
      float t_64 = 0.f;
      if (c.n64 > 0) {
        std::vector<float> t;
        for (int r = 0; r <= reps && rc == DFT_OK; ++r) {
          (void)hipEventRecord(e0, s);
          hipError_t e = launch_combo(p, p->ws, ci, false, 0, c.n64, c.shape64, P, Pa, Pb, s);
          (void)hipEventRecord(e1, s);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          if (e != hipSuccess) rc = fail(DFT_EHIP, std::string("tune: ") + hipGetErrorString(e));
          if (r > 0) t.push_back(ms);
        }
        if (rc == DFT_OK) {
          std::sort(t.begin(), t.end());
          t_64 = t[t.size() / 2];
        }
      }
      c.use32 = t_64 + t_32 < 0.97f * t_all;
    }
  }
  // whole builds with the class launches serial and concurrent (FP64 mode,
  // whole-plan tuning only); keep the faster
  if (rc == DFT_OK && combo < 0 && !p->aux.empty()) {
    float tbest[2] = {1e30f, 1e30f};
    for (int conc = 0; conc < 2 && rc == DFT_OK; ++conc) {
      p->concurrent = conc == 1;
      std::vector<double*> tmp(3, nullptr);
      for (int m = 0; m < 1 + p->nk && rc == DFT_OK; ++m)
        if (hipMalloc(&tmp[m], std::max<long long>(p->nelem, 1) * sizeof(double)) != hipSuccess)
          rc = fail(DFT_EHIP, "tune: hipMalloc");
      for (int r = 0; r <= reps && rc == DFT_OK; ++r) {
        (void)hipEventRecord(e0, s);
        rc = jk_enqueue(p, p->ws, P, Pa, Pb, tmp[0], tmp[1], tmp[2], DFT_FP64, s, false);
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (rc == DFT_OK && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            r > 0)
          tbest[conc] = std::min(tbest[conc], ms);
      }
      for (double* x : tmp) (void)hipFree(x);
    }
    p->concurrent = tbest[1] < tbest[0];
    // mixed mode: the per-class FP32 choices above come from launches timed
    // alone; keep them only if the whole mixed build is faster with them
    // than with FP32 in every class
    std::vector<char> chosen;
    bool any64 = false;
    for (const Combo& c : p->combos) {
      chosen.push_back(c.use32);
      any64 = any64 || (!c.use32 && c.q.size() > (size_t)c.n64);
    }
    if (any64 && rc == DFT_OK) {
      std::vector<double*> tmp(3, nullptr);
      for (int m = 0; m < 1 + p->nk && rc == DFT_OK; ++m)
        if (hipMalloc(&tmp[m], std::max<long long>(p->nelem, 1) * sizeof(double)) != hipSuccess)
          rc = fail(DFT_EHIP, "tune: hipMalloc");
      float tm[2] = {1e30f, 1e30f};
      for (int all32 = 0; all32 < 2 && rc == DFT_OK; ++all32) {
        for (size_t ci = 0; ci < p->combos.size(); ++ci) p->combos[ci].use32 = all32 ? true : (bool)chosen[ci];
        for (int r = 0; r <= reps && rc == DFT_OK; ++r) {
          (void)hipEventRecord(e0, s);
          rc = jk_enqueue(p, p->ws, P, Pa, Pb, tmp[0], tmp[1], tmp[2], DFT_MIXED, s, false);
          (void)hipEventRecord(e1, s);
          float ms = 0.f;
          if (rc == DFT_OK && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
              r > 0)
            tm[all32] = std::min(tm[all32], ms);
        }
      }
      for (double* x : tmp) (void)hipFree(x);
      const bool keep = tm[0] <= tm[1];
      for (size_t ci = 0; ci < p->combos.size(); ++ci) p->combos[ci].use32 = keep ? (bool)chosen[ci] : true;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  p->tuned = true;
  if (report)
    for (size_t ci = 0; ci < p->combos.size() && (int)(6 * ci + 5) < report_cap; ++ci) {
      const Combo& c = p->combos[ci];
      const int v[6] = {c.mfma ? 1 : 0, c.shape64.block, c.shape64.cap, c.use32 ? 0 : 1, c.shape32.block,
                        c.shape32.cap};
      for (int k = 0; k < 6; ++k) report[6 * ci + k] = v[k];
    }
  return DFT_OK;
}

int dft_plan_set_shapes(dft_plan p, const int* shapes, int n) {
  if (!p || !shapes) return fail(DFT_EINVAL, "NULL argument");
  // two passes: every shape is checked before any is applied
  for (int apply = 0; apply < 2; ++apply) {
    for (size_t ci = 0; ci < p->combos.size() && (int)(6 * ci + 5) < n; ++ci) {
      Combo& c = p->combos[ci];
      Shape s64, s32;
      s64.wave = shapes[6 * ci];
      s64.block = shapes[6 * ci + 1];
      s64.cap = shapes[6 * ci + 2];
      s32.wave = shapes[6 * ci + 3];
      s32.block = shapes[6 * ci + 4];
      s32.cap = shapes[6 * ci + 5];
      if (s64.wave != 0 && !(s64.wave == 1 && mfma_combo((int)ci)))
        return fail(DFT_EINVAL, "the first int of the FP64 shape must be 0, or 1 (matrix cores) for a heavy combination");
      const bool use_mfma = s64.wave == 1;
      s64.wave = 0;
      if (s32.wave != 0 && s32.wave != 1) return fail(DFT_EINVAL, "the first int of the FP32 shape must be 0 or 1");
      const bool use32 = s32.wave == 0;
      s32.wave = 0;
      for (const Shape* s : {&s64, &s32}) {
        if (s->block != 64 && s->block != 128 && s->block != 256 && s->block != 512)
          return fail(DFT_EINVAL, "block must be 64, 128, 256 or 512");
        if (s->block > block_max((int)ci)) return fail(DFT_EINVAL, "block exceeds the class's largest workgroup");
        if (s->cap < 1) return fail(DFT_EINVAL, "grid cap must be positive");
      }
      if (!apply) continue;
      c.shape64 = s64;
      c.shape32 = s32;
      c.use32 = use32;
      c.mfma = use_mfma;
    }
  }
  return DFT_OK;
}

int dft_plan_profile(dft_plan p, int nsteps) {
  if (!p || nsteps < 0) return fail(DFT_EINVAL, "bad arguments");
  for (hipEvent_t e : p->prof_ev) (void)hipEventDestroy(e);
  p->prof_ev.clear();
  p->prof_steps = p->prof_used = 0;
  if (nsteps == 0) return DFT_OK;
  const int nslot = 2 * (int)p->combos.size() + 2;
  p->prof_ev.resize((size_t)nsteps * nslot * 2);
  for (auto& e : p->prof_ev) HIP_TRY(hipEventCreate(&e));
  p->prof_steps = nsteps;
  return DFT_OK;
}

int dft_plan_profile_read(dft_plan p, double* ms_sum, int cap, int* nslot_out, int* nsteps_out) {
  if (!p) return fail(DFT_EINVAL, "plan is NULL");
  const int nslot = 2 * (int)p->combos.size() + 2;
  if (nslot_out) *nslot_out = nslot;
  if (nsteps_out) *nsteps_out = p->prof_used;
  if (!ms_sum) return DFT_OK;
  for (int k = 0; k < std::min(cap, nslot); ++k) ms_sum[k] = 0.0;
  for (int st = 0; st < p->prof_used; ++st)
    for (int k = 0; k < std::min(cap, nslot); ++k) {
      hipEvent_t* ev = &p->prof_ev[((size_t)st * nslot + k) * 2];
      HIP_TRY(hipEventSynchronize(ev[1]));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
      ms_sum[k] += ms;
    }
  return DFT_OK;
}

int dft_eri_quartets(dft_plan p, const int32_t* quartets, int64_t nq, double* out) {
  if (!p) return fail(DFT_EINVAL, "plan is NULL");
  if (nq < 0 || (nq > 0 && (!quartets || !out))) return fail(DFT_EINVAL, "bad quartet list");
  const int n = p->nsh;
  struct Req {
    int2 q;
    int perm[4];  // internal position -> requested position
    long long out_off;
  };
  std::vector<std::vector<Req>> by_combo(p->combos.size());
  long long off = 0;
  for (int64_t r = 0; r < nq; ++r) {
    int s[4], ii[4];
    for (int k = 0; k < 4; ++k) {
      s[k] = quartets[4 * r + k];
      if (s[k] < 0 || s[k] >= n) return fail(DFT_EINVAL, "quartet shell index out of range");
      ii[k] = p->in2int[s[k]];
    }
    if (p->sh[ii[0]].mol != p->sh[ii[1]].mol || p->sh[ii[0]].mol != p->sh[ii[2]].mol ||
        p->sh[ii[0]].mol != p->sh[ii[3]].mol)
      return fail(DFT_EINVAL, "quartet shells belong to different molecules");
    // positions of the internal bra (A >= B) and ket (C >= D)
    int pos[4] = {0, 1, 2, 3};
    if (ii[0] < ii[1]) std::swap(pos[0], pos[1]);
    if (ii[2] < ii[3]) std::swap(pos[2], pos[3]);
    int pb = p->pair_of.at(pair_key(ii[pos[0]], ii[pos[1]]));
    int pk = p->pair_of.at(pair_key(ii[pos[2]], ii[pos[3]]));
    if (p->pair_pc[pb] < p->pair_pc[pk]) {
      std::swap(pb, pk);
      std::swap(pos[0], pos[2]);
      std::swap(pos[1], pos[3]);
    }
    Req q;
    q.q = make_int2(pb, pk);
    for (int k = 0; k < 4; ++k) q.perm[k] = pos[k];
    q.out_off = off;
    int sz = 1;
    for (int k = 0; k < 4; ++k) sz *= ncart_h(p->sh[ii[k]].l);
    off += sz;
    by_combo[combo_index(p->pair_pc[pb], p->pair_pc[pk])].push_back(q);
  }
  for (size_t ci = 0; ci < by_combo.size(); ++ci) {
    const auto& reqs = by_combo[ci];
    if (reqs.empty()) continue;
    int la, lb, lc, ld;
    pc_moments(p->combos[ci].pcb, la, lb);
    pc_moments(p->combos[ci].pck, lc, ld);
    const int nI[4] = {ncart_h(la), ncart_h(lb), ncart_h(lc), ncart_h(ld)};
    const int blk = nI[0] * nI[1] * nI[2] * nI[3];
    std::vector<int2> q(reqs.size());
    for (size_t k = 0; k < reqs.size(); ++k) q[k] = reqs[k].q;
    std::vector<double> raw;
    int rc = raw_blocks(p, (int)ci, q, raw, blk);
    if (rc) return rc;
    for (size_t k = 0; k < reqs.size(); ++k) {
      const Req& R = reqs[k];
      int nreq[4];
      for (int i = 0; i < 4; ++i) nreq[R.perm[i]] = nI[i];
      int x[4];
      for (x[0] = 0; x[0] < nI[0]; ++x[0])
        for (x[1] = 0; x[1] < nI[1]; ++x[1])
          for (x[2] = 0; x[2] < nI[2]; ++x[2])
            for (x[3] = 0; x[3] < nI[3]; ++x[3]) {
              int y[4];
              for (int i = 0; i < 4; ++i) y[R.perm[i]] = x[i];
              const long long o = ((long long)(y[0] * nreq[1] + y[1]) * nreq[2] + y[2]) * nreq[3] + y[3];
              out[R.out_off + o] = raw[k * blk + ((x[0] * nI[1] + x[1]) * nI[2] + x[2]) * nI[3] + x[3]];
            }
    }
  }
  return DFT_OK;
}

int dft_oneelec(dft_plan p, double* S, double* T, double* V, void* stream) {
  if (!p || !S || !T || !V) return fail(DFT_EINVAL, "NULL argument");
  hipError_t e = hipSuccess;
  for (int pc = 0; pc < NPC && e == hipSuccess; ++pc) {
    if (p->shp_n[pc] == 0) continue;
    int la, lb;
    pc_moments(pc, la, lb);
    e = launch_one_electron(la, lb, p->d_shpairs + p->shp_first[pc], p->shp_n[pc], p->d_sh, p->d_exps, p->d_coefs,
                            p->d_atoms, S, T, V, p->d_boys, (hipStream_t)stream);
  }
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("one-electron: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_ao_grid(dft_plan p, const float* grid_xyz, int64_t npts, double* ao, int64_t ld, void* stream) {
  if (!p || !grid_xyz || !ao) return fail(DFT_EINVAL, "NULL argument");
  if (p->nmol != 1) return fail(DFT_EUNSUPPORTED, "grid values take a single-molecule plan");
  if (npts < 0 || ld < npts) return fail(DFT_EINVAL, "need 0 <= npts <= ld");
  hipError_t e = launch_ao_grid(p->nsh, p->d_sh, p->d_exps, p->d_coefs, grid_xyz, (long long)npts, ao, (long long)ld,
                                (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("ao grid: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_grad(dft_plan p, const int32_t* shell_atom, const double* P, const double* Pa, const double* Pb,
             const double* W, double* grad, void* stream) {
  if (!p) return fail(DFT_EINVAL, "plan is NULL");
  if (!shell_atom || !P || !Pa || !W || !grad) return fail(DFT_EINVAL, "NULL argument");
  if (p->natom <= 0) return fail(DFT_EINVAL, "plan has no nuclei");
  if (p->nmol != 1) return fail(DFT_EUNSUPPORTED, "gradients take a single-molecule plan");
  if (!p->grad_ready) {
    int rc = grad_prepare(p, shell_atom);
    if (rc) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  const int nbf = p->mol_nbf[0];
  for (size_t ci = 0; ci < p->combos.size(); ++ci) {
    const Combo& c = p->combos[ci];
    const long long nq = p->gq_n[ci];
    if (nq == 0) continue;
    int la, lb, lc, ld;
    pc_moments(c.pcb, la, lb);
    pc_moments(c.pck, lc, ld);
    hipError_t e = launch_eri_grad(la, lb, lc, ld, p->d_gq[ci], p->d_gslots_ab[ci], nq, p->d_gpairs, p->d_grec,
                                   p->d_shell_ao, P, Pa, Pb, nbf, p->d_gstage, p->d_boys, s);
    if (e == hipSuccess)
      e = launch_eri_grad(lc, ld, la, lb, p->d_gq_swap[ci], p->d_gslots_cd[ci], nq, p->d_gpairs, p->d_grec,
                          p->d_shell_ao, P, Pa, Pb, nbf, p->d_gstage, p->d_boys, s);
    if (e != hipSuccess) return fail(DFT_EHIP, std::string("eri_grad: ") + hipGetErrorString(e));
  }
  hipError_t e = hipSuccess;
  for (int pc = 0; pc < NPC && e == hipSuccess; ++pc) {
    if (p->g1_n[pc] == 0) continue;
    int la, lb;
    pc_moments(pc, la, lb);
    e = launch_grad1e(la, lb, p->d_g1sp, p->d_g1pidx + p->g1_first[pc], p->g1_n[pc], p->d_sh, p->d_exps, p->d_coefs,
                      p->natom, p->d_atoms, P, W, nbf, p->d_g1slots, p->d_gstage, p->d_boys, s);
  }
  if (e == hipSuccess)
    e = launch_grad_reduce(p->natom, p->n_gchunk, p->d_gchunks, p->d_gcfirst, p->d_gbase, p->d_gstage, p->d_gpart,
                           grad, s);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("grad 1e/reduce: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_nuclear_prim(int nfa, int nfb, const double* exp_a, const double* exp_b, const double* cen_a,
                     int cen_a_stride, const double* cen_b, int cen_b_stride, const int32_t* pow_a,
                     const int32_t* pow_b, int natom, const int32_t* atoms, const double* atom_xyz, double* V,
                     double* Ex, double* Ey, double* Ez, double* p, double* P, int32_t* nx, int32_t* ny, int32_t* nz,
                     void* stream) {
  if (nfa < 0 || nfb < 0 || natom < 0) return fail(DFT_EINVAL, "negative size");
  if ((cen_a_stride != 0 && cen_a_stride != 3) || (cen_b_stride != 0 && cen_b_stride != 3))
    return fail(DFT_EINVAL, "centre stride must be 0 (one centre) or 3 (per function)");
  double* tab = nullptr;
  int rc = boys_table_device(&tab);
  if (rc) return rc;
  hipError_t e = launch_nuclear_prim(nfa, nfb, exp_a, exp_b, cen_a, cen_a_stride, cen_b, cen_b_stride, pow_a, pow_b,
                                     natom, atoms, atom_xyz, V, Ex, Ey, Ez, p, P, nx, ny, nz, tab, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("nuclear_prim: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_eri_prim(int64_t n_ab, const double* p_ab, const double* P_ab, const double* Ex_ab, const double* Ey_ab,
                 const double* Ez_ab, const int32_t* nx_ab, const int32_t* ny_ab, const int32_t* nz_ab, int64_t n_cd,
                 const double* p_cd, const double* P_cd, const double* Ex_cd, const double* Ey_cd,
                 const double* Ez_cd, const int32_t* nx_cd, const int32_t* ny_cd, const int32_t* nz_cd, int e_stride,
                 double* out, void* stream) {
  if (n_ab < 0 || n_cd < 0 || e_stride < 1) return fail(DFT_EINVAL, "bad sizes");
  double* tab = nullptr;
  int rc = boys_table_device(&tab);
  if (rc) return rc;
  hipError_t e = launch_eri_prim(n_ab, p_ab, P_ab, Ex_ab, Ey_ab, Ez_ab, nx_ab, ny_ab, nz_ab, n_cd, p_cd, P_cd, Ex_cd,
                                 Ey_cd, Ez_cd, nx_cd, ny_cd, nz_cd, e_stride, out, tab, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("eri_prim: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_syevj_tol(int n, int batch, double* A, double* w, double* V, int init_v, int32_t* sweeps, double tol,
                  void* stream) {
  return dft_syevj_ws(n, batch, A, w, V, init_v, sweeps, tol, nullptr, 0, stream);
}

int64_t dft_syevj_ws_bytes(int n) {
  if (n <= 0 || n > JAC_NMAX) return 0;
  return (int64_t)syevj_log_bytes(n);
}

int dft_syevj_ws(int n, int batch, double* A, double* w, double* V, int init_v, int32_t* sweeps, double tol,
                 void* work, int64_t work_bytes, void* stream) {
  if (n <= 0 || batch < 0) return fail(DFT_EINVAL, "n must be positive and batch >= 0");
  if (n > JAC_NMAX) return fail(DFT_EUNSUPPORTED, "dft_syevj: order " + std::to_string(n) + " above " +
                                                      std::to_string(JAC_NMAX));
  if (!(tol >= 0.0)) return fail(DFT_EINVAL, "dft_syevj_tol: tol must be >= 0");
  if (batch > 0 && (!A || !w || !V)) return fail(DFT_EINVAL, "NULL argument");
  const long long per = syevj_log_bytes(n);
  if (work && work_bytes < per * (long long)batch)
    return fail(DFT_EINVAL, "dft_syevj_ws: work_bytes " + std::to_string(work_bytes) + " below batch x " +
                                std::to_string(per));
  const long long nn = (long long)n * n;
  hipError_t e = launch_syevj(n, batch, A, nn, w, n, V, nn, init_v ? 1 : 0, sweeps, tol, (char*)work,
                              work ? per : 0, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("syevj: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_syevj(int n, int batch, double* A, double* w, double* V, int init_v, int32_t* sweeps, void* stream) {
  return dft_syevj_tol(n, batch, A, w, V, init_v, sweeps, 0.0, stream);
}

int dft_scf_diis_stage(int n, int history, const double* F, const double* R, double* hist_F, double* hist_R,
                       double* G, const double* P, const double* H, const double* Pa, double* state, double* Fu,
                       void* stream) {
  static_assert(DFT_SCF_HMAX == SCF_HMAX && DFT_SCF_ST_SLOT == SCF_ST_SLOT && DFT_SCF_ST_COEF == SCF_ST_COEF &&
                    DFT_SCF_STATE_LEN == SCF_STATE_LEN,
                "state layout");
  if (n <= 0 || n > JAC_NMAX) return fail(DFT_EUNSUPPORTED, "dft_scf_diis_stage: order must be in [1, 128]");
  if (history < 1 || history > DFT_SCF_HMAX) return fail(DFT_EINVAL, "history must be in [1, 16]");
  if (!F || !R || !hist_F || !hist_R || !G || !P || !H || !Pa || !state || !Fu) return fail(DFT_EINVAL, "NULL argument");
  hipError_t e = launch_scf_diis_stage(n, history, F, R, hist_F, hist_R, G, P, H, Pa, state, Fu, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("scf diis stage: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_scf_damp_control(int n, double* Pa, const double* Pauf, double* P, double* state, void* stream) {
  if (n <= 0 || n > JAC_NMAX) return fail(DFT_EUNSUPPORTED, "dft_scf_damp_control: order must be in [1, 128]");
  if (!Pa || !Pauf || !P || !state) return fail(DFT_EINVAL, "NULL argument");
  hipError_t e = launch_scf_damp_control(n, Pa, Pauf, P, state, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("scf damp/control: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_scf_gemm(int m, int ncol, int k, const double* A, int lda, int trans_a, const double* B, int ldb,
                 int trans_b, double* C, int ldc, const double* copy_src, double* copy_dst, void* stream) {
  if (m < 1 || ncol < 1 || k < 1 || m > JAC_NMAX || ncol > JAC_NMAX || k > JAC_NMAX)
    return fail(DFT_EUNSUPPORTED, "dft_scf_gemm: m, ncol, k must be in [1, 128]");
  if (!A || !B || !C || (copy_src == nullptr) != (copy_dst == nullptr)) return fail(DFT_EINVAL, "NULL argument");
  if (lda < (trans_a ? m : k) || ldb < (trans_b ? k : ncol) || ldc < ncol)
    return fail(DFT_EINVAL, "dft_scf_gemm: leading dimension below the row length");
  hipError_t e = launch_scf_gemm(m, ncol, k, A, lda, trans_a ? 1 : 0, B, ldb, trans_b ? 1 : 0, C, ldc, copy_src,
                                 copy_dst, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("scf gemm: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_scf_fock_residual(int n, const double* H, const double* J, const double* K, const double* PS, double* F,
                          double* R, void* stream) {
  if (n < 1 || n > JAC_NMAX) return fail(DFT_EUNSUPPORTED, "dft_scf_fock_residual: order must be in [1, 128]");
  if (!H || !J || !K || !PS || !F || !R) return fail(DFT_EINVAL, "NULL argument");
  hipError_t e = launch_scf_fock_residual(n, H, J, K, PS, F, R, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("scf fock/residual: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_scf_orbital_matrix(int n, const double* XQ, const double* T, double* A, const double* Q, double* V,
                           void* stream) {
  if (n < 1 || n > JAC_NMAX) return fail(DFT_EUNSUPPORTED, "dft_scf_orbital_matrix: order must be in [1, 128]");
  if (!XQ || !T || !A || !Q || !V) return fail(DFT_EINVAL, "NULL argument");
  hipError_t e = launch_scf_orbital_matrix(n, XQ, T, A, Q, V, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("scf orbital matrix: ") + hipGetErrorString(e));
  return DFT_OK;
}

int dft_boys(const double* T, int n, int mmax, double* out, int fp32, void* stream) {
  if (n < 0 || (n > 0 && (!T || !out))) return fail(DFT_EINVAL, "bad arguments");
  if (mmax < 0 || mmax > 16) return fail(DFT_EUNSUPPORTED, "mmax must be in [0, 16]");
  double* tab = nullptr;
  int rc = boys_table_device(&tab);
  if (rc) return rc;
  hipError_t e = launch_boys(T, n, mmax, out, fp32, tab, (hipStream_t)stream);
  if (e != hipSuccess) return fail(DFT_EHIP, std::string("boys: ") + hipGetErrorString(e));
  return DFT_OK;
}

}  // extern "C"
