// Fused ERI + J/K digest kernels and raw-block kernels for s/p class
// combinations, instantiated per (LA LB | LC LD) and per precision.
#include "eri_quartet.h"
#include "kernels.h"

namespace dft {

constexpr int ERI_BLOCK = 256;  // raw kernel

// Largest J/K workgroup per class: 512 lanes leave 128 VGPRs per lane, which
// the classes up to (pp|ps) fit; the larger ones keep 256 (and 256 VGPRs).
constexpr int jk_block_max(int la, int lb, int lc, int ld) {
  return ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) <= 27 ? 512 : 256;
}

// Blocks of one kind from the 64 lanes of a wave: lanes whose block goes to
// the same target form runs (mask bit l set where lane l starts a run); each
// run is summed across its lanes by a segmented inclusive scan (six shift
// steps, a fixed order, so the sums are reproducible) and its last lane
// stores the sum at its slot (slot < 0: the lane stores nothing). A mask of
// only singleton runs stores directly.
template <int NE>
__device__ __forceinline__ void store_runs(double (&v)[NE], unsigned long long mask, int lane, long long slot,
                                           double* __restrict__ stage) {
  if (mask != ~0ull) {
    const unsigned long long upto = mask & (~0ull >> (63 - lane));
    const int start = 63 - __builtin_clzll(upto);  // bit 0 is always set
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const bool take = lane - d >= start;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const double u = __shfl_up(v[e], d, 64);
        if (take) v[e] += u;
      }
    }
  }
  if (slot >= 0) {
#pragma unroll
    for (int e = 0; e < NE; ++e) stage[slot + e] = v[e];
  }
}

// Each lane evaluates one canonical quartet (bra pair >= ket pair in the
// class-sorted pair order) and forms its contribution blocks
//   Jab, Jcd, then per exchange density Kac, Kad, Kbc, Kbd;
// the blocks of a wave's lanes that share a target are summed in the wave
// (store_runs) and stored once at stage + slots[k * ntot + q] + ch * (block
// size) for primitive chunk ch (block-major slot table, so that each slot
// load of a wave is one contiguous run; a run's chunks own consecutive
// blocks of its target). A split combination with a lane table puts the bra
// chunks of a quartet in neighbouring lanes, so its runs also sum over bra
// chunks; only the ket chunks keep blocks of their own.
// With w = 2^-([a=b]+[c=d]+[ab=cd]):
//   Jab = 2w sum_cd I P_cd        Jcd = 2w sum_ab I P_ab
//   Kac = w sum_bd I P_bd  Kad = w sum_bc I P_bc  Kbc = w sum_ad I P_ad  Kbd = w sum_ac I P_ac
// and the half matrices Jh, Kh summed over all blocks give J = Jh + Jh^T and
// K = Kh + Kh^T, i.e. the eight orientations and degeneracy weights of
// fill_matrices (shelled.py:791-859). No atomics: the slots of one target
// (ordered shell pair of a half matrix) are contiguous, assigned in quartet
// order at plan time, and summed by a separate fixed-order pass. The quartet
// lists are padded so that a wave's quartets (and hence its run sums) are
// the same in every launch split.
template <int LA, int LB, int LC, int LD, typename real, int NK>
__global__ __launch_bounds__(jk_block_max(LA, LB, LC, LD)) void eri_jk_kernel(
    const int2* __restrict__ quartets, const long long* __restrict__ slots,
    const unsigned long long* __restrict__ segmask, int first, int n, int ntot, int splb, int splk,
    const int4* __restrict__ lanes, const int4* __restrict__ pairs, const real* __restrict__ recb, int sb, int fb,
    const real* __restrict__ reck, const int4* __restrict__ shell_dm, const double* __restrict__ DJ,
    const double* __restrict__ DK0, const double* __restrict__ DK1, double* __restrict__ stage,
    const double* __restrict__ boys_gtab) {
  using S = QuartetShape<LA, LB, LC, LD, real>;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND;
  constexpr int NBLK = 2 + 4 * NK;
  __shared__ double btab[BOYS_LDS];
  boys_stage<S::L>(boys_gtab, btab);
  // Without a lane table, lane v: chunk ch = v / n (chunk-major, so the
  // lanes of a wave share the chunk and mostly the ket pair), quartet
  // first + v % n. With one, lane v: ket chunk ck = v / n, table entry
  // first + v % n (the bra chunks of a quartet are neighbouring entries).
  const unsigned nv = (unsigned)n * (unsigned)(lanes ? splk : splb * splk);
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
    const unsigned ch = v / (unsigned)n;
    // idx: this lane's position in the slot table and the run masks
    const int idx = first + (int)(v - ch * (unsigned)n);
    int q = idx, ck = (int)ch, tb0 = 0, tbs = 1;
    bool idle = false;
    if (lanes) {
      const int4 t = lanes[idx];
      q = t.x;
      tb0 = t.y;
      tbs = max(1, t.z);
      idle = t.w != 0;
    } else {
      ck = (int)ch % splk;
    }
    const int2 qd = quartets[q];
    // padding quartets (bra -1) take their ket as bra and evaluate nothing;
    // idle table entries repeat a quartet of their wave and evaluate nothing
    const bool live = qd.x >= 0 && !idle;
    const int4 pb = pairs[qd.x >= 0 ? qd.x : qd.y];
    const int4 pk = pairs[qd.y];
    // this lane's primitive pairs of the bra pair (nbp of them from i0,
    // bstride apart) and of the ket pair ([j0, j0 + nkp)); a table entry
    // takes bra primitives tb0, tb0 + tbs, ... below the pair's count in the
    // record set of this launch
    const int szk = (pk.w + splk - 1) / splk;
    const int j0 = min(pk.w, ck * szk), nkp = min(pk.w, j0 + szk) - j0;
    int i0, nbp, bstride = 1;
    if (lanes) {
      i0 = min(pb.w, tb0);
      nbp = live ? (pb.w - i0 + tbs - 1) / tbs : 0;
      bstride = tbs;
    } else {
      const int cb = (int)ch / splk, szb = (pb.w + splb - 1) / splb;
      i0 = min(pb.w, cb * szb);
      nbp = live ? min(pb.w, i0 + szb) - i0 : 0;
    }
    const unsigned ob = (unsigned)((qd.x >= 0 ? qd.x : qd.y) - fb) + (unsigned)(i0 * S::RSB * sb);
    const unsigned ok = (unsigned)pk.z + (unsigned)(j0 * S::RSK);
    real I[S::NI];
    // When the whole wave shares the ket primitive range (the common case in
    // ket-major order) its offset and length are made explicitly
    // wave-uniform, so the inner-loop record loads become scalar loads and
    // the loop a scalar loop; otherwise the same code runs per lane.
    const unsigned okf = (unsigned)__builtin_amdgcn_readfirstlane((int)ok);
    const int nkf = __builtin_amdgcn_readfirstlane(nkp);
    // the primitive-quartet cut pays in FP64 (the FP32 records are already
    // cut per pair, and the test costs a larger share of the FP32 loop)
    constexpr bool cut = std::is_same<real, double>::value;
    if (__all(ok == okf && nkp == nkf)) {
      quartet_integrals<LA, LB, LC, LD, real, cut>(recb, sb, ob, nbp, reck, 1, okf, nkf, btab, I, PRIM_QCUT64,
                                                   bstride);
    } else {
      quartet_integrals<LA, LB, LC, LD, real, cut>(recb, sb, ob, nbp, reck, 1, ok, nkp, btab, I, PRIM_QCUT64,
                                                   bstride);
    }
    double w = 1.0;
    if (pb.x == pb.y) w *= 0.5;
    if (pk.x == pk.y) w *= 0.5;
    if (qd.x == qd.y) w *= 0.5;
    // density / J / K of the quartet's molecule: packed matrix at mb, stride nbf
    const int4 sa = shell_dm[pb.x];
    const int a0 = sa.x, b0 = shell_dm[pb.y].x;
    const int c0 = shell_dm[pk.x].x, d0 = shell_dm[pk.y].x;
    const int nbf = sa.y;
    const long long mb = (long long)(unsigned)sa.z | ((long long)sa.w << 32);
    const double* DJm = DJ + mb;
    const long long sstr = ntot;
    const long long* sl = slots + idx;
    // slot of block kind k for this lane's chunk (the chunks of a run own
    // consecutive blocks of ne elements; with a lane table the bra chunks
    // are summed in the wave and only the ket chunk has blocks of its own)
    const int srow = lanes ? ck : (int)ch;
    auto at = [&](long long s, int ne) { return s >= 0 ? s + (long long)srow * ne : -1LL; };
    // the wave's index in the combination's quartet list (or lane table),
    // the same for its 64 lanes (n and first are multiples of the wave, and
    // so its chunk)
    const int lane = threadIdx.x & 63;
    const unsigned wid = (unsigned)__builtin_amdgcn_readfirstlane(idx / WAVE_QUARTETS);
    const unsigned long long* msk = segmask + (size_t)wid * NBLK;
    // Coulomb
    {
      double Pcd[S::NCD], Pab[S::NAB];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) Pcd[c * ND + d] = DJm[(long long)(c0 + c) * nbf + d0 + d];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) Pab[a * NB + b] = DJm[(long long)(a0 + a) * nbf + b0 + b];
      double jab[S::NAB], jcd[S::NCD];
#pragma unroll
      for (int cd = 0; cd < S::NCD; ++cd) jcd[cd] = 0.0;
#pragma unroll
      for (int ab = 0; ab < S::NAB; ++ab) {
        double s = 0.0;
#pragma unroll
        for (int cd = 0; cd < S::NCD; ++cd) {
          const double v = (double)I[ab * S::NCD + cd];
          s = fma(v, Pcd[cd], s);
          jcd[cd] = fma(v, Pab[ab], jcd[cd]);
        }
        jab[ab] = 2.0 * w * s;
      }
#pragma unroll
      for (int cd = 0; cd < S::NCD; ++cd) jcd[cd] *= 2.0 * w;
      store_runs(jab, msk[0], lane, at(sl[0], S::NAB), stage);
      store_runs(jcd, msk[1], lane, at(sl[sstr], S::NCD), stage);
    }
    // Exchange, one density at a time. The densities are symmetric, and
    // P_bd is read as P_db (likewise P_bc, P_ad, P_ac): the lanes of a wave
    // mostly share the ket shells c, d, so their reads fall in the rows of
    // c and d instead of one row per lane.
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const double* D = (k == 0 ? DK0 : DK1) + mb;
      double Pbd[NB * ND], Pbc[NB * NC], Pad[NA * ND], Pac[NA * NC];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int d = 0; d < ND; ++d) Pbd[b * ND + d] = D[(long long)(d0 + d) * nbf + b0 + b];
#pragma unroll
        for (int c = 0; c < NC; ++c) Pbc[b * NC + c] = D[(long long)(c0 + c) * nbf + b0 + b];
      }
#pragma unroll
      for (int a = 0; a < NA; ++a) {
#pragma unroll
        for (int d = 0; d < ND; ++d) Pad[a * ND + d] = D[(long long)(d0 + d) * nbf + a0 + a];
#pragma unroll
        for (int c = 0; c < NC; ++c) Pac[a * NC + c] = D[(long long)(c0 + c) * nbf + a0 + a];
      }
      double kac[NA * NC], kad[NA * ND], kbc[NB * NC], kbd[NB * ND];
#pragma unroll
      for (int x = 0; x < NA * NC; ++x) kac[x] = 0.0;
#pragma unroll
      for (int x = 0; x < NA * ND; ++x) kad[x] = 0.0;
#pragma unroll
      for (int x = 0; x < NB * NC; ++x) kbc[x] = 0.0;
#pragma unroll
      for (int x = 0; x < NB * ND; ++x) kbd[x] = 0.0;
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
              const double v = (double)I[((a * NB + b) * NC + c) * ND + d];
              kac[a * NC + c] = fma(v, Pbd[b * ND + d], kac[a * NC + c]);
              kad[a * ND + d] = fma(v, Pbc[b * NC + c], kad[a * ND + d]);
              kbc[b * NC + c] = fma(v, Pad[a * ND + d], kbc[b * NC + c]);
              kbd[b * ND + d] = fma(v, Pac[a * NC + c], kbd[b * ND + d]);
            }
#pragma unroll
      for (int x = 0; x < NA * NC; ++x) kac[x] *= w;
#pragma unroll
      for (int x = 0; x < NA * ND; ++x) kad[x] *= w;
#pragma unroll
      for (int x = 0; x < NB * NC; ++x) kbc[x] *= w;
#pragma unroll
      for (int x = 0; x < NB * ND; ++x) kbd[x] *= w;
      const long long* ks = sl + (2 + 4 * k) * sstr;
      const unsigned long long* km = msk + 2 + 4 * k;
      store_runs(kac, km[0], lane, at(ks[0], NA * NC), stage);
      store_runs(kad, km[1], lane, at(ks[sstr], NA * ND), stage);
      store_runs(kbc, km[2], lane, at(ks[2 * sstr], NB * NC), stage);
      store_runs(kbd, km[3], lane, at(ks[3 * sstr], NB * ND), stage);
    }
  }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(ERI_BLOCK) void eri_raw_kernel(const int2* __restrict__ quartets, long long nq,
                                                            const int4* __restrict__ pairs,
                                                            const double* __restrict__ rec,
                                                            double* __restrict__ out,
                                                            const double* __restrict__ boys_gtab) {
  using S = QuartetShape<LA, LB, LC, LD, double>;
  __shared__ double btab[BOYS_LDS];
  boys_stage<S::L>(boys_gtab, btab);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += stride) {
    const int2 qd = quartets[q];
    const int4 pb = pairs[qd.x];
    const int4 pk = pairs[qd.y];
    double I[S::NI];
    quartet_integrals<LA, LB, LC, LD, double>(rec, 1, (unsigned)pb.z, pb.w, rec, 1, (unsigned)pk.z, pk.w, btab,
                                              I);
    double* o = out + q * S::NI;
#pragma unroll
    for (int x = 0; x < S::NI; ++x) o[x] = I[x];
  }
}

static int grid_for(long long nq, int cap, int block = ERI_BLOCK) {
  long long g = (nq + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

template <int LA, int LB, int LC, int LD, typename real>
static hipError_t jk_one(const int2* q, const long long* sl, const unsigned long long* sm, const int4* pairs,
                         const RecView& rv,
                         const int4* ao, const double* DJ, const double* DK0, const double* DK1, int nk,
                         double* stage, const double* bt, int blk, int cap, hipStream_t s) {
  if (rv.n <= 0) return hipSuccess;
  // whole waves only: every wave runs its shuffles with all 64 lanes
  if (rv.n % WAVE_QUARTETS || rv.first % WAVE_QUARTETS || rv.ntot % WAVE_QUARTETS) return hipErrorInvalidValue;
  const long long nq = (long long)rv.n * (rv.lanes ? 1 : rv.splb) * rv.splk;
  if (blk != 64 && blk != 128 && blk != 256 && blk != 512) return hipErrorInvalidValue;
  if (blk > jk_block_max(LA, LB, LC, LD)) return hipErrorInvalidValue;
  dim3 grid(grid_for(nq, cap, blk)), block(blk);
  if (nk == 1)
    hipLaunchKernelGGL((eri_jk_kernel<LA, LB, LC, LD, real, 1>), grid, block, 0, s, q, sl, sm, rv.first, rv.n,
                       rv.ntot, rv.splb, rv.splk, rv.lanes, pairs, (const real*)rv.bra, rv.sb, rv.fb, (const real*)rv.ket,
                       ao, DJ, DK0, DK1, stage, bt);
  else
    hipLaunchKernelGGL((eri_jk_kernel<LA, LB, LC, LD, real, 2>), grid, block, 0, s, q, sl, sm, rv.first, rv.n,
                       rv.ntot, rv.splb, rv.splk, rv.lanes, pairs, (const real*)rv.bra, rv.sb, rv.fb, (const real*)rv.ket,
                       ao, DJ, DK0, DK1, stage, bt);
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
static hipError_t raw_one(const int2* q, long long nq, const int4* pairs, const double* rec, double* out,
                          const double* bt, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  dim3 grid(grid_for(nq, 1 << 20)), block(ERI_BLOCK);
  hipLaunchKernelGGL((eri_raw_kernel<LA, LB, LC, LD>), grid, block, 0, s, q, nq, pairs, rec, out, bt);
  return hipGetLastError();
}

// combo index = pcb (pcb+1)/2 + pck over pair classes ss=0, ps=1, pp=2,
// ds=3, dp=4, dd=5; the lane kernel is built for the s/p combinations and
// the d combinations whose integrals, G and R fit a lane's registers
#define DFT_FOR_COMBOS(X) \
  X(0, 0, 0, 0, 0)        \
  X(1, 1, 0, 0, 0)        \
  X(2, 1, 0, 1, 0)        \
  X(3, 1, 1, 0, 0)        \
  X(4, 1, 1, 1, 0)        \
  X(5, 1, 1, 1, 1)        \
  X(6, 2, 0, 0, 0)        \
  X(7, 2, 0, 1, 0)        \
  X(9, 2, 0, 2, 0)        \
  X(10, 2, 1, 0, 0)       \
  X(15, 2, 2, 0, 0)

int lane_block_max(int combo) {
  switch (combo) {
#define X(id, la, lb, lc, ld) \
  case id:                    \
    return jk_block_max(la, lb, lc, ld);
    DFT_FOR_COMBOS(X)
#undef X
    default:
      return 0;
  }
}

bool lane_kernel_combo(int combo) {
  switch (combo) {
#define X(id, la, lb, lc, ld) \
  case id:                    \
    return true;
    DFT_FOR_COMBOS(X)
#undef X
    default:
      return false;
  }
}

hipError_t launch_eri_jk(int combo, bool fp32, const int2* q, const long long* sl, const unsigned long long* sm,
                         const int4* pairs, const RecView& rv, const int4* ao, const double* DJ, const double* DK0,
                         const double* DK1, int nk, double* stage, const double* bt, int block, int cap,
                         hipStream_t s) {
  switch (combo) {
#define X(id, la, lb, lc, ld)                                                                          \
  case id:                                                                                             \
    return fp32 ? jk_one<la, lb, lc, ld, float>(q, sl, sm, pairs, rv, ao, DJ, DK0, DK1, nk, stage, bt, \
                                                 block, cap, s)                                            \
                : jk_one<la, lb, lc, ld, double>(q, sl, sm, pairs, rv, ao, DJ, DK0, DK1, nk, stage, bt, \
                                                  block, cap, s);
    DFT_FOR_COMBOS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_eri_raw(int combo, const int2* q, long long nq, const int4* pairs, const double* rec,
                          double* out, const double* bt, hipStream_t s) {
  switch (combo) {
#define X(id, la, lb, lc, ld) \
  case id:                    \
    return raw_one<la, lb, lc, ld>(q, nq, pairs, rec, out, bt, s);
    DFT_FOR_COMBOS(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace dft
