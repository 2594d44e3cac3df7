// Contracted shell-quartet ERIs, one quartet per lane, for s/p shells.
//
// For a quartet of class (LA LB | LC LD) the lane loops over primitive bra
// pairs i and ket pairs j (records built by the pair pre-pass), evaluates
// F_m and R^0_tuv once per primitive quartet and contracts in two steps:
//   ket step  G[h][cd] += pref * sum_{tau nu phi} (-1)^{tau+nu+phi}
//                          E^cd_tau E^cd_nu E^cd_phi R[h + (tau,nu,phi)]
//   bra step  I[ab][cd] += sum_h E^ab_t E^ab_u E^ab_v G[h][cd]
// with the bra step applied once per bra primitive pair (G is summed over
// ket primitives first). pref = 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd, with the
// 1/p, 1/q factors stored in the pair records.
//
// Reference: eri_kernel (shelled.py:452-532) computes the same integral per
// primitive Cartesian quartet as C_x C_y C_z . R (convolution + box sum) and
// contracts afterwards with einsum (shelled.py:663-689). Contraction
// coefficients, primitive normalisation and the exp(-mu R_AB^2) factor are
// folded into K_ab; the angular normalisation 1/sqrt((2l-1)!!) per axis is
// folded into the E tables, so I[ab][cd] is directly the normalised,
// contracted integral of shelled.py:668-689.
#pragma once
#include "md_device.h"

namespace dft {

// Pair-record access. Element f of primitive pair i of one shell pair sits
// at base[f * s + off + i * RS * s]:
//   plain layout        s = 1, off = the pair's record offset;
//   interleaved layout  base = the pair class's block, s = pairs in the class,
//                       off = the pair's index within the class,
// so in the interleaved layout the lanes of a wave working on neighbouring
// pairs of a class read neighbouring addresses (coalesced), and base + f*s
// is wave-uniform (scalar base register, 32-bit vector offset).
// Byte offsets are formed in 32 bits (record blocks stay below 4 GiB), so
// the load takes the scalar-base + 32-bit vector-offset form.
template <typename real>
__device__ __forceinline__ real fld(const real* __restrict__ base, int s, int f, unsigned off) {
  const char* __restrict__ b = reinterpret_cast<const char*>(base + f * s);
  return *reinterpret_cast<const real*>(b + off * (unsigned)sizeof(real));
}

// E_t^{ia,ib} of one axis (first element AX of that axis' table in the
// record); E_0^{00} = 1 by construction (the exp(-mu X_AB^2) factor lives in
// K_ab), so that entry is never loaded.
template <int LB, int IA, int IB, int T, int AX, typename real>
__device__ __forceinline__ real e_at(const real* __restrict__ base, int s, unsigned off) {
  if constexpr (IA == 0 && IB == 0) return real(1);
  else return fld(base, s, AX + eoff(LB, IA, IB) + T, off);
}

constexpr double TWO_PI_2P5 = 34.986836655249725693;  // 2 pi^{5/2}

template <int LA, int LB, int LC, int LD, typename real>
struct QuartetShape {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NAB = NA * NB, NCD = NC * ND, NI = NAB * NCD;
  static constexpr int L = LA + LB + LC + LD;
  static constexpr int LAB = LA + LB;
  static constexpr int NHB = nherm(LAB);
  static constexpr int NEB = ne_axis(LA, LB), NEK = ne_axis(LC, LD);
  static constexpr int RSB = rec_size(LA, LB), RSK = rec_size(LC, LD);
};

// Number of ket component pairs cd evaluated by one call: all of them
// (NCDK < 0) or the slice [CD0, CD0 + NCDK).
template <int LA, int LB, int LC, int LD, int NCDK>
constexpr int ncd_slice() { return NCDK < 0 ? ncart(LC) * ncart(LD) : NCDK; }

// Primitive-quartet cut of the J/K kernels: a primitive quartet whose
// Schwarz bound s_ab s_cd (record heads) is below it is skipped. At most K^4
// such terms (K^4 <= 1296 here) change an integral by < 1.3e-11 (FP64) or
// 1.3e-10 (FP32, whose integrals carry larger errors anyway).
constexpr double PRIM_QCUT64 = 1e-14;
constexpr float PRIM_QCUT32 = 1e-13f;

// I[ab * NK + k] = (ab|cd) for cd = CD0 + k, k < NK (NK = ncd_slice), for the
// quartet (bra pair: base bb, stride sb, offset ob, nb primitive pairs; ket
// pair likewise). With the default slice this is the whole I[ab * NCD + cd].
// With CUT, primitive quartets with s_ab s_cd < qcut are left out: the
// primitive pairs of a pair are stored in decreasing s, so each lane stops
// its bra loop at the first bra primitive below the cut with the largest ket
// factor, and the ket loop ends once no lane of the wave is above it (the
// terms a lane leaves out are exactly its own sub-cut ones). The bra loop
// visits nb primitive pairs bstride apart (a lane of a split quartet takes
// every bstride-th one, still in decreasing s).
template <int LA, int LB, int LC, int LD, typename real, bool CUT = false, int CD0 = 0, int NCDK = -1>
__device__ __forceinline__ void quartet_integrals(
    const real* __restrict__ bb, int sb, unsigned ob, int nb, const real* __restrict__ bk, int sk, unsigned ok,
    int nk, const double* __restrict__ btab,
    real (&I)[QuartetShape<LA, LB, LC, LD, real>::NAB * ncd_slice<LA, LB, LC, LD, NCDK>()], real qcut = real(0),
    int bstride = 1) {
  using S = QuartetShape<LA, LB, LC, LD, real>;
  constexpr int NK = ncd_slice<LA, LB, LC, LD, NCDK>();
  constexpr int EXB = REC_HEAD, EYB = REC_HEAD + S::NEB, EZB = REC_HEAD + 2 * S::NEB;
  constexpr int EXK = REC_HEAD, EYK = REC_HEAD + S::NEK, EZK = REC_HEAD + 2 * S::NEK;
#pragma unroll
  for (int x = 0; x < S::NAB * NK; ++x) I[x] = real(0);
  const real sk0 = CUT && nk > 0 ? fld(bk, sk, 5, ok) : real(0);  // largest ket factor
  for (int i = 0; i < nb; ++i) {
    const unsigned oi = ob + (unsigned)(i * bstride * S::RSB * sb);
    const real sab = CUT ? fld(bb, sb, 5, oi) : real(0);
    if (CUT && sab * sk0 < qcut) break;
    const real p = fld(bb, sb, 0, oi), Px = fld(bb, sb, 1, oi), Py = fld(bb, sb, 2, oi);
    const real Pz = fld(bb, sb, 3, oi), Kab = fld(bb, sb, 4, oi);
    real G[S::NHB * NK];
#pragma unroll
    for (int x = 0; x < S::NHB * NK; ++x) G[x] = real(0);
    // the head (q, Q, K_cd) of ket primitive j + 1 is loaded while primitive
    // j is evaluated, which hides most of the L2 latency of the record loads
    unsigned oj = ok;
    real nq = real(0), nQx = real(0), nQy = real(0), nQz = real(0), nK = real(0), nS = real(0);
    if (nk > 0) {
      nq = fld(bk, sk, 0, oj), nQx = fld(bk, sk, 1, oj), nQy = fld(bk, sk, 2, oj);
      nQz = fld(bk, sk, 3, oj), nK = fld(bk, sk, 4, oj);
      if constexpr (CUT) nS = fld(bk, sk, 5, oj);
    }
    for (int j = 0; j < nk; ++j) {
      const real q = nq, X = Px - nQx, Y = Py - nQy, Z = Pz - nQz, Kcd = nK;
      const bool on = !CUT || sab * nS >= qcut;
      if (CUT && !__any(on)) break;
      const unsigned ojc = oj;
      if (j + 1 < nk) oj += (unsigned)(S::RSK * sk);
      nq = fld(bk, sk, 0, oj), nQx = fld(bk, sk, 1, oj), nQy = fld(bk, sk, 2, oj);
      nQz = fld(bk, sk, 3, oj), nK = fld(bk, sk, 4, oj);
      if constexpr (CUT) nS = fld(bk, sk, 5, oj);
      // K_ab and K_cd carry 1/p and 1/q (pair pre-pass), so
      // pref = 2 pi^{5/2} K_ab K_cd / sqrt(p+q) and alpha = p q / (p+q)
      const real rs = rsqrt_nr(p + q);
      const real alpha = p * q * rs * rs;
      // a lane's sub-cut terms are multiplied by exactly 0 (the wave runs on
      // while any lane is above the cut)
      const real pref = (!CUT || on) ? real(TWO_PI_2P5) * Kab * Kcd * rs : real(0);
      const real T = alpha * (X * X + Y * Y + Z * Z);
      real F[S::L + 1];
      boys_eval<S::L, real>(T, btab, F);
      real R[nherm(S::L)];
      hermite_R<S::L, real>(F, real(-2) * alpha, X, Y, Z, R);
      static_for<0, NK>([&](auto kI) {
        constexpr int kk = decltype(kI)::value;
        constexpr int cd = CD0 + kk;
        constexpr int c = cd / S::ND, d = cd % S::ND;
        constexpr int cx = comp_x(LC, c), cy = comp_y(LC, c), cz = comp_z(LC, c);
        constexpr int dx = comp_x(LD, d), dy = comp_y(LD, d), dz = comp_z(LD, d);
        // E^cd products with the (-1)^{tau+nu+phi} sign and pref folded in
        real Ek[(cx + dx + 1) * (cy + dy + 1) * (cz + dz + 1)];
        static_for<0, cx + dx + 1>([&](auto tI) {
          static_for<0, cy + dy + 1>([&](auto uI) {
            static_for<0, cz + dz + 1>([&](auto vI) {
              constexpr int ta = decltype(tI)::value, nu = decltype(uI)::value, ph = decltype(vI)::value;
              constexpr int k = (ta * (cy + dy + 1) + nu) * (cz + dz + 1) + ph;
              real e = e_at<LD, cx, dx, ta, EXK>(bk, sk, ojc) * e_at<LD, cy, dy, nu, EYK>(bk, sk, ojc) *
                       e_at<LD, cz, dz, ph, EZK>(bk, sk, ojc) * pref;
              Ek[k] = ((ta + nu + ph) & 1) ? -e : e;
            });
          });
        });
        static_for<0, S::NHB>([&](auto hI) {
          constexpr int h = decltype(hI)::value;
          constexpr int t = h_t(h), u = h_u(h), v = h_v(h);
          real s = G[h * NK + kk];
          static_for<0, cx + dx + 1>([&](auto tI) {
            static_for<0, cy + dy + 1>([&](auto uI) {
              static_for<0, cz + dz + 1>([&](auto vI) {
                constexpr int ta = decltype(tI)::value, nu = decltype(uI)::value, ph = decltype(vI)::value;
                constexpr int k = (ta * (cy + dy + 1) + nu) * (cz + dz + 1) + ph;
                s = fma(Ek[k], R[hidx(t + ta, u + nu, v + ph)], s);
              });
            });
          });
          G[h * NK + kk] = s;
        });
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
            constexpr int h = hidx(t, u, v);
            const real e = e_at<LB, ax, bx, t, EXB>(bb, sb, oi) * e_at<LB, ay, by, u, EYB>(bb, sb, oi) *
                           e_at<LB, az, bz, v, EZB>(bb, sb, oi);
            static_for<0, NK>([&](auto kI) {
              constexpr int kk = decltype(kI)::value;
              I[ab * NK + kk] = fma(e, G[h * NK + kk], I[ab * NK + kk]);
            });
          });
        });
      });
    });
  }
}

}  // namespace dft
