// ERIs of the heavy class combinations (a d shell with p/d partners), one
// lane per (quartet, ket component pair cd), and their J/K digest.
//
// A whole (dp|dp) or (dd|dd) quartet does not fit a lane's registers
// ((dd|dd): 1296 integrals, 165 Hermite R values), so lane (q, cd) computes
// the column I[ab][cd] only: per primitive quartet it evaluates Boys and R
// (recomputed per cd, which costs less than the contraction it serves), the
// ket step G[h] = pref sum_{tau nu phi} (-1)^{tau+nu+phi} E^c_tau E^d_nu E_phi
// R[h + (tau,nu,phi)] over the dense ket Hermite set with the lane's E
// factors loaded at run time (zero beyond its component sums), and, once per
// bra primitive, the bra step for all ab with compile-time indices
// (eri_quartet.h does the same with every index compile-time). cd is a
// run-time index so that one kernel serves all columns of a class (a
// compile-time column per instantiation would multiply the code by NCD).
// The column goes to out[q * NI + ab * NCD + cd]: the raw integral blocks
// (parity tests, Schwarz diagonals) or a per-combination buffer from which
// a second kernel, one wave per quartet, forms the same J/K contribution
// blocks as the lane kernel (kernels_eri.hip).
#include <algorithm>

#include "eri_quartet.h"
#include "kernels.h"

namespace dft {

// I[ab] = (ab|cd) for one run-time ket column cd.
template <int LA, int LB, int LC, int LD, typename real>
__device__ __forceinline__ void quartet_column(const real* __restrict__ bb, int sb, unsigned ob, int nb,
                                               const real* __restrict__ bk, unsigned ok, int nk, int cd,
                                               const double* __restrict__ btab,
                                               real (&I)[ncart(LA) * ncart(LB)], real qcut) {
  using S = QuartetShape<LA, LB, LC, LD, real>;
  constexpr int LCD = LC + LD;
  constexpr int EXB = REC_HEAD, EYB = REC_HEAD + S::NEB, EZB = REC_HEAD + 2 * S::NEB;
  // component sums and E offsets of the column's (c, d) per axis (-1 marks
  // E_0^{00} = 1, which is not stored); scalar work when cd is wave-uniform
  const int c = cd / S::ND, d = cd - (cd / S::ND) * S::ND;
  int exa[3], eoa[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int ic = comp_of(LC, c, a), id = comp_of(LD, d, a);
    exa[a] = ic + id;
    eoa[a] = (ic == 0 && id == 0) ? -1 : eoff(LD, ic, id);
  }
  const int exx = exa[0], exy = exa[1], exz = exa[2];
  const int eox = eoa[0], eoy = eoa[1], eoz = eoa[2];
#pragma unroll
  for (int x = 0; x < S::NAB; ++x) I[x] = real(0);
  // primitive-quartet cut as in quartet_integrals (records in decreasing
  // Schwarz factor; a lane's sub-cut terms are left out exactly; qcut = 0:
  // none)
  const real sk0 = nk > 0 ? bk[ok + 5] : real(0);
  for (int i = 0; i < nb; ++i) {
    const unsigned oi = ob + (unsigned)(i * S::RSB * sb);
    const real sab = fld(bb, sb, 5, oi);
    if (sab * sk0 < qcut) break;
    const real p = fld(bb, sb, 0, oi), Px = fld(bb, sb, 1, oi), Py = fld(bb, sb, 2, oi);
    const real Pz = fld(bb, sb, 3, oi), Kab = fld(bb, sb, 4, oi);
    real G[S::NHB];
#pragma unroll
    for (int x = 0; x < S::NHB; ++x) G[x] = real(0);
    for (int j = 0; j < nk; ++j) {
      const real* rk = bk + ok + (unsigned)(j * S::RSK);
      const bool on = sab * rk[5] >= qcut;
      if (!__any(on)) break;
      const real q = rk[0];
      const real X = Px - rk[1], Y = Py - rk[2], Z = Pz - rk[3];
      const real rs = rsqrt_nr(p + q);
      const real alpha = p * q * rs * rs;
      const real pref = on ? real(TWO_PI_2P5) * Kab * rk[4] * rs : real(0);
      const real T = alpha * (X * X + Y * Y + Z * Z);
      real F[S::L + 1];
      boys_eval<S::L, real>(T, btab, F);
      real R[nherm(S::L)];
      hermite_R<S::L, real>(F, real(-2) * alpha, X, Y, Z, R);
      // the lane's ket E factors with the (-1)^t sign and pref (x axis)
      // folded in; zero beyond the component sums
      real ex[LCD + 1], ey[LCD + 1], ez[LCD + 1];
      const real* E = rk + REC_HEAD;
#pragma unroll
      for (int t = 0; t <= LCD; ++t) {
        const real sg = (t & 1) ? real(-1) : real(1);
        ex[t] = t > exx ? real(0) : (eox < 0 ? real(1) : E[eox + t]) * sg * pref;
        ey[t] = t > exy ? real(0) : (eoy < 0 ? real(1) : E[S::NEK + eoy + t]) * sg;
        ez[t] = t > exz ? real(0) : (eoz < 0 ? real(1) : E[2 * S::NEK + eoz + t]) * sg;
      }
      // G[h] += sum_tau ex[tau] sum_nu ey[nu] sum_phi ez[phi] R[h + (tau,nu,phi)],
      // nested so that no ey*ez table is held (registers are what limits
      // these lanes)
      static_for<0, S::NHB>([&](auto hI) {
        constexpr int h = decltype(hI)::value;
        constexpr int t = h_t(h), u = h_u(h), v = h_v(h);
        real g = G[h];
        static_for<0, LCD + 1>([&](auto tI) {
          constexpr int ta = decltype(tI)::value;
          real inner = real(0);
          static_for<0, LCD + 1 - ta>([&](auto nI) {
            constexpr int nu = decltype(nI)::value;
            real in2 = real(0);
            static_for<0, LCD + 1 - ta - nu>([&](auto pI) {
              constexpr int ph = decltype(pI)::value;
              in2 = fma(ez[ph], R[hidx(t + ta, u + nu, v + ph)], in2);
            });
            inner = fma(ey[nu], in2, inner);
          });
          g = fma(ex[ta], inner, g);
        });
        G[h] = g;
      });
    }
    static_for<0, S::NAB>([&](auto abI) {
      constexpr int ab = decltype(abI)::value;
      constexpr int a = ab / S::NB, b = ab % S::NB;
      constexpr int ax = comp_x(LA, a), ay = comp_y(LA, a), az = comp_z(LA, a);
      constexpr int bx = comp_x(LB, b), by = comp_y(LB, b), bz = comp_z(LB, b);
      static_for<0, ax + bx + 1>([&](auto tI) {
        static_for<0, ay + by + 1>([&](auto uI) {
          static_for<0, az + bz + 1>([&](auto vI) {
            constexpr int t = decltype(tI)::value, u = decltype(uI)::value, v = decltype(vI)::value;
            const real e = e_at<LB, ax, bx, t, EXB>(bb, sb, oi) * e_at<LB, ay, by, u, EYB>(bb, sb, oi) *
                           e_at<LB, az, bz, v, EZB>(bb, sb, oi);
            I[ab] = fma(e, G[hidx(t, u, v)], I[ab]);
          });
        });
      });
    });
  }
}

template <int LA, int LB, int LC, int LD, typename real>
__global__ __launch_bounds__(256) void eri_chunk_kernel(const int2* __restrict__ quartets, int first, int n,
                                                        int splb, const int4* __restrict__ pairs,
                                                        const real* __restrict__ recb, int sb, int fb,
                                                        const real* __restrict__ reck, double* __restrict__ out,
                                                        long long xs, long long qs, long long ss,
                                                        const double* __restrict__ boys_gtab, real qcut) {
  using S = QuartetShape<LA, LB, LC, LD, real>;
  __shared__ double btab[BOYS_LDS];
  boys_stage<S::L>(boys_gtab, btab);
  // lane v: bra primitive chunk sp, column cd, quartet first + qi with
  // v = (sp * NCD + cd) * n + qi (a wave shares sp and cd, and mostly the
  // ket pair); chunk sp's partial integrals go to out + sp * ss
  const unsigned ncol = (unsigned)n * (unsigned)S::NCD;
  const unsigned nv = ncol * (unsigned)splb;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
    const int sp = (int)(v / ncol);
    const unsigned rem = v - (unsigned)sp * ncol;
    const int cd = (int)(rem / (unsigned)n);
    const int q = first + (int)(rem - (unsigned)cd * (unsigned)n);
    const int2 qd = quartets[q];
    const int4 pb = pairs[qd.x];
    const int4 pk = pairs[qd.y];
    const int szb = (pb.w + splb - 1) / splb;
    const int i0 = min(pb.w, sp * szb), nbp = min(pb.w, i0 + szb) - i0;
    const unsigned ob = (unsigned)(qd.x - fb) + (unsigned)(i0 * S::RSB * sb);
    real I[S::NAB];
    // wave-uniform ket and column (the common case): scalar ket loads and
    // table lookups, as in the lane kernel
    const unsigned okf = (unsigned)__builtin_amdgcn_readfirstlane(pk.z);
    const int nkf = __builtin_amdgcn_readfirstlane(pk.w);
    const int cdf = __builtin_amdgcn_readfirstlane(cd);
    if (__all((unsigned)pk.z == okf && pk.w == nkf && cd == cdf))
      quartet_column<LA, LB, LC, LD, real>(recb, sb, ob, nbp, reck, okf, nkf, cdf, btab, I, qcut);
    else
      quartet_column<LA, LB, LC, LD, real>(recb, sb, ob, nbp, reck, (unsigned)pk.z, pk.w, cd, btab, I, qcut);
    double* o = out + (long long)sp * ss + (long long)q * qs;
#pragma unroll
    for (int ab = 0; ab < S::NAB; ++ab) o[(long long)(ab * S::NCD + cd) * xs] = (double)I[ab];
  }
}

// J/K blocks of one quartet from its integrals I[q * NI + x] (x = abcd
// row-major), one 64-lane wave per quartet, lanes over the output elements
// of the quartet's blocks (run-time loops: the blocks of the heavy classes
// are large and their digest is a small part of the work); same weights and
// blocks as the lane kernel:
//   Jab = 2w sum_cd I P_cd        Jcd = 2w sum_ab I P_ab
//   Kac = w sum_bd I P_bd  Kad = w sum_bc I P_bc  Kbc = w sum_ad I P_ad  Kbd = w sum_ac I P_ac
constexpr int DIGEST_WAVES = 4;       // quartets (waves) per workgroup
constexpr int DIGEST_NI_MAX = 1296;   // integrals of a (dd|dd) quartet
constexpr int DIGEST_BLK_MAX = 36;    // elements of a (d, d) density block

template <int NK>
__global__ __launch_bounds__(256) void jk_digest_kernel(int na, int nb, int nc, int nd, int splb, long long ss,
                                                        const int2* __restrict__ quartets,
                                                        const long long* __restrict__ slots, int n,
                                                        const int4* __restrict__ pairs,
                                                        const int4* __restrict__ shell_dm,
                                                        const double* __restrict__ DJ, const double* __restrict__ DK0,
                                                        const double* __restrict__ DK1,
                                                        const double* __restrict__ ibuf, double* __restrict__ stage) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int nab = na * nb, ncd = nc * nd, ni = nab * ncd;
  const int nK = na * nc + na * nd + nb * nc + nb * nd;
  const int nout = nab + ncd + NK * nK;
  // the wave's quartet integrals, summed over the bra primitive chunks in a
  // fixed order, and the six density blocks the quartet contracts with
  // (P_cd, P_ab, then per exchange density P_bd, P_bc, P_ad, P_ac) are
  // staged in LDS with coalesced loads; the output loops below then read LDS
  // only (the density reads were one global load per multiply-add)
  __shared__ double sI[DIGEST_WAVES][DIGEST_NI_MAX];
  __shared__ double sP[DIGEST_WAVES][2 + 4 * NK][DIGEST_BLK_MAX];
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double* Iw = sI[wv];
  double* Pw = &sP[wv][0][0];
  for (int q = blockIdx.x * wpb + wv; q < n; q += gridDim.x * wpb) {
    const int2 qd = quartets[q];
    const int4 pb = pairs[qd.x];
    const int4 pk = pairs[qd.y];
    double w = 1.0;
    if (pb.x == pb.y) w *= 0.5;
    if (pk.x == pk.y) w *= 0.5;
    if (qd.x == qd.y) w *= 0.5;
    const int4 sa = shell_dm[pb.x];
    const long long a0 = sa.x, b0 = shell_dm[pb.y].x, c0 = shell_dm[pk.x].x, d0 = shell_dm[pk.y].x;
    const long long nbf = sa.y;
    const long long mb = (long long)(unsigned)sa.z | ((long long)sa.w << 32);
    const double* Iq = ibuf + (long long)q * ni;
    __builtin_amdgcn_wave_barrier();  // the previous quartet's reads are done
    for (int x = lane; x < ni; x += 64) {
      double v = Iq[x];
      for (int sp = 1; sp < splb; ++sp) v += Iq[sp * ss + x];
      Iw[x] = v;
    }
    // block t: rows r0.. (nr of them) and columns k0.. (nc_ of them) of D
    auto stage_block = [&](int t, const double* D, long long r0, int nr, long long k0, int nc_) {
      for (int x = lane; x < nr * nc_; x += 64) {
        const int r = x / nc_, k = x - r * nc_;
        Pw[t * DIGEST_BLK_MAX + x] = D[(r0 + r) * nbf + k0 + k];
      }
    };
    stage_block(0, DJ + mb, c0, nc, d0, nd);
    stage_block(1, DJ + mb, a0, na, b0, nb);
    for (int kk = 0; kk < NK; ++kk) {
      const double* D = (kk == 0 ? DK0 : DK1) + mb;
      stage_block(2 + 4 * kk, D, b0, nb, d0, nd);
      stage_block(3 + 4 * kk, D, b0, nb, c0, nc);
      stage_block(4 + 4 * kk, D, a0, na, d0, nd);
      stage_block(5 + 4 * kk, D, a0, na, c0, nc);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Every output element is a sum over two indices, s = sum_i sum_j
    // I[ib + i * i1 + j * i2] * P[pb + i * p1 + j * p2] (in the order the
    // block formulas above list them), so one loop body serves all blocks
    // and a wave's lanes do not diverge over the block kinds.
    const long long* sl = slots + (long long)q * (2 + 4 * NK);
    for (int o = lane; o < nout; o += 64) {
      int ib, i1, i2, n1, n2, pb_, p1, p2, blk, x;
      double scale;
      if (o < nab) {
        const int a = o / nb, b = o - a * nb;
        ib = (a * nb + b) * ncd, i1 = nd, i2 = 1, n1 = nc, n2 = nd;
        pb_ = 0, p1 = nd, p2 = 1, blk = 0, x = o, scale = 2.0 * w;
      } else if (o < nab + ncd) {
        x = o - nab;
        const int c = x / nd, d = x - c * nd;
        ib = c * nd + d, i1 = nb * ncd, i2 = ncd, n1 = na, n2 = nb;
        pb_ = DIGEST_BLK_MAX, p1 = nb, p2 = 1, blk = 1, scale = 2.0 * w;
      } else {
        x = o - nab - ncd;
        const int kk = x / nK;
        x -= kk * nK;
        const int t0 = 2 + 4 * kk;
        scale = w;
        if (x < na * nc) {
          const int a = x / nc, c = x - a * nc;  // K_ac: sum over b, d of I P_bd
          ib = a * nb * ncd + c * nd, i1 = ncd, i2 = 1, n1 = nb, n2 = nd;
          pb_ = t0 * DIGEST_BLK_MAX, p1 = nd, p2 = 1, blk = t0;
        } else if ((x -= na * nc) < na * nd) {
          const int a = x / nd, d = x - a * nd;  // K_ad: sum over b, c of I P_bc
          ib = a * nb * ncd + d, i1 = ncd, i2 = nd, n1 = nb, n2 = nc;
          pb_ = (t0 + 1) * DIGEST_BLK_MAX, p1 = nc, p2 = 1, blk = t0 + 1;
        } else if ((x -= na * nd) < nb * nc) {
          const int b = x / nc, c = x - b * nc;  // K_bc: sum over a, d of I P_ad
          ib = b * ncd + c * nd, i1 = nb * ncd, i2 = 1, n1 = na, n2 = nd;
          pb_ = (t0 + 2) * DIGEST_BLK_MAX, p1 = nd, p2 = 1, blk = t0 + 2;
        } else {
          x -= nb * nc;
          const int b = x / nd, d = x - b * nd;  // K_bd: sum over a, c of I P_ac
          ib = b * ncd + d, i1 = nb * ncd, i2 = nd, n1 = na, n2 = nc;
          pb_ = (t0 + 3) * DIGEST_BLK_MAX, p1 = nc, p2 = 1, blk = t0 + 3;
        }
      }
      double s = 0.0;
      for (int i = 0; i < n1; ++i) {
        const double* Ir = Iw + ib + i * i1;
        const double* Pr = Pw + pb_ + i * p1;
        for (int j = 0; j < n2; ++j) s = fma(Ir[j * i2], Pr[j * p2], s);
      }
      stage[sl[blk] + x] = scale * s;
    }
  }
}

static int chunk_grid(long long lanes, int cap, int block) {
  long long g = (lanes + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

// heavy combinations: (ds|pp), (dp|ps), (dp|pp), (dp|ds), (dp|dp), (dd|ps),
// (dd|pp), (dd|ds), (dd|dp), (dd|dd)
#define DFT_FOR_HEAVY(X) \
  X(8, 2, 0, 1, 1)       \
  X(11, 2, 1, 1, 0)      \
  X(12, 2, 1, 1, 1)      \
  X(13, 2, 1, 2, 0)      \
  X(14, 2, 1, 2, 1)      \
  X(16, 2, 2, 1, 0)      \
  X(17, 2, 2, 1, 1)      \
  X(18, 2, 2, 2, 0)      \
  X(19, 2, 2, 2, 1)      \
  X(20, 2, 2, 2, 2)

int chunk_slices(int combo) {
  switch (combo) {
#define X(id, la, lb, lc, ld) \
  case id:                    \
    return ncart(lc) * ncart(ld);
    DFT_FOR_HEAVY(X)
#undef X
    default:
      return 0;
  }
}

hipError_t launch_eri_chunk(int combo, bool fp32, const int2* q, int first, int n, int splb, const int4* pairs,
                            const RecView& rv, double* out, long long xs, long long qs, long long ss,
                            const double* bt, int block, int cap, bool cut, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
  switch (combo) {
#define X(id, la, lb, lc, ld)                                                                                   \
  case id: {                                                                                                    \
    const long long lanes = (long long)n * chunk_slices(id) * splb;                                             \
    dim3 grid(chunk_grid(lanes, cap, block)), blk(block);                                                       \
    if (fp32)                                                                                                   \
      hipLaunchKernelGGL((eri_chunk_kernel<la, lb, lc, ld, float>), grid, blk, 0, s, q, first, n, splb, pairs,   \
                         (const float*)rv.bra, rv.sb, rv.fb, (const float*)rv.ket, out, xs, qs, ss, bt,        \
                         cut ? PRIM_QCUT32 : 0.0f);                                                             \
    else                                                                                                        \
      hipLaunchKernelGGL((eri_chunk_kernel<la, lb, lc, ld, double>), grid, blk, 0, s, q, first, n, splb, pairs,  \
                         (const double*)rv.bra, rv.sb, rv.fb, (const double*)rv.ket, out, xs, qs, ss, bt,      \
                         cut ? PRIM_QCUT64 : 0.0);                                                              \
    return hipGetLastError();                                                                                   \
  }
    DFT_FOR_HEAVY(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_jk_digest(int combo, const int2* q, const long long* sl, int n, int splb, long long ss,
                            const int4* pairs, const int4* shell_dm, const double* DJ, const double* DK0,
                            const double* DK1, int nk, const double* ibuf, double* stage, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int la, lb, lc, ld;
  switch (combo) {
#define X(id, a, b, c, d) \
  case id:                \
    la = a;               \
    lb = b;               \
    lc = c;               \
    ld = d;               \
    break;
    DFT_FOR_HEAVY(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  const int wpb = DIGEST_WAVES;  // quartets (waves) per workgroup (the LDS staging is sized for it)
  if (ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > DIGEST_NI_MAX) return hipErrorInvalidValue;
  if (ncart(2) * ncart(2) > DIGEST_BLK_MAX || std::max(std::max(la, lb), std::max(lc, ld)) > 2)
    return hipErrorInvalidValue;  // density blocks of at most (d, d)
  dim3 grid((unsigned)std::min<long long>(((long long)n + wpb - 1) / wpb, 65536)), blk(64 * wpb);
  if (nk == 1)
    hipLaunchKernelGGL((jk_digest_kernel<1>), grid, blk, 0, s, ncart(la), ncart(lb), ncart(lc), ncart(ld), splb,
                       ss, q, sl, n, pairs, shell_dm, DJ, DK0, DK1, ibuf, stage);
  else
    hipLaunchKernelGGL((jk_digest_kernel<2>), grid, blk, 0, s, ncart(la), ncart(lb), ncart(lc), ncart(ld), splb,
                       ss, q, sl, n, pairs, shell_dm, DJ, DK0, DK1, ibuf, stage);
  return hipGetLastError();
}

}  // namespace dft
