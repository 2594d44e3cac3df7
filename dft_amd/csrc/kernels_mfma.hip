// Heavy class combinations on the matrix cores: the two McMurchie-Davidson
// contractions of a shell quartet as FP64 MFMA products.
//
// For one primitive quartet the ket step is the small dense product
//   G[h][cd] += R~[h][k] E~[k][cd],   R~[h][k] = R[h + k] (Hermite sums),
//   E~[k][cd] = (-1)^{tau+nu+phi} pref E^cd_tau E^cd_nu E^cd_phi
// over the bra Hermite set h (t+u+v <= la+lb) and the ket Hermite set
// k = (tau, nu, phi) (<= lc+ld), and once per bra primitive the bra step
//   I[ab][cd] += E~[ab][h] G[h][cd]
// (E~ zero outside the component sums of ab / cd). These are the
// convolution and box sum of the reference's eri_kernel
// (shelled.py:213-226,475-531), contracted over primitives as in calc_eri
// (shelled.py:663-689).
//
// One wave per (quartet, bra primitive chunk); v_mfma_f64_16x16x4f64 with
// its gfx950 lane layout (lane l: A[l%16][l/16], B[l/16][l%16], result item
// r = D[4r + l/16][l%16]). The result layout of a G tile is the B-operand
// layout of the bra product's K steps, so G passes from one product to the
// next in registers. R^0_tuv is built level by level in LDS (three entries
// per lane) and gathered into the A operands; the E~ entries are formed by
// each lane from the pair records. Padding rows and columns are zero.
#include <hip/hip_runtime.h>

#include "eri_quartet.h"
#include "kernels.h"

namespace dft {

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_d4 mfma16x16x4(double a, double b, mfma_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Packed E~ entry of one lane: offsets of the three axis factors in the
// record's E tables (+1; 0 = the unstored E_0^{00} = 1), the sign, validity.
__device__ __forceinline__ int pack_e(int ox, int oy, int oz, int neg, int valid) {
  return (ox + 1) | (oy + 1) << 6 | (oz + 1) << 12 | neg << 18 | valid << 19;
}
__device__ __forceinline__ double eval_e(int pk, const double* __restrict__ e, int ne) {
  if (!(pk >> 19)) return 0.0;
  const int ox = (pk & 63) - 1, oy = ((pk >> 6) & 63) - 1, oz = ((pk >> 12) & 63) - 1;
  const double x = ox < 0 ? 1.0 : e[ox];
  const double y = oy < 0 ? 1.0 : e[ne + oy];
  const double z = oz < 0 ? 1.0 : e[2 * ne + oz];
  const double v = x * y * z;
  return (pk >> 18) & 1 ? -v : v;
}

// E~ entry for Cartesian pair component index cc of a pair class (l1, l2)
// and Hermite index hh (bra: no sign; ket: (-1)^{tau+nu+phi}).
__device__ __forceinline__ int e_entry(int l1, int l2, int cc, int hh, int nh, int ncc, bool ket) {
  if (cc >= ncc || hh >= nh) return 0;
  const int n2 = (l2 + 1) * (l2 + 2) / 2;
  const int c1 = cc / n2, c2 = cc - c1 * n2;
  const int t = h_t(hh), u = h_u(hh), v = h_v(hh);
  const int hv[3] = {t, u, v};
  int off[3];
  for (int a = 0; a < 3; ++a) {
    const int i1 = comp_of(l1, c1, a), i2 = comp_of(l2, c2, a);
    if (hv[a] > i1 + i2) return 0;
    off[a] = (i1 == 0 && i2 == 0) ? -1 : eoff(l2, i1, i2) + hv[a];
  }
  return pack_e(off[0], off[1], off[2], ket ? ((t + u + v) & 1) : 0, 1);
}

template <int LA, int LB, int LC, int LD>
struct MfmaShape {
  static constexpr int L = LA + LB + LC + LD;
  static constexpr int NR = nherm(L);                    // R^0 entries
  static constexpr int NHB = nherm(LA + LB), NHK = nherm(LC + LD);
  static constexpr int NAB = ncart(LA) * ncart(LB), NCD = ncart(LC) * ncart(LD);
  static constexpr int MTG = (NHB + 15) / 16;            // G row tiles (bra Hermite)
  static constexpr int KTG = (NHK + 3) / 4;              // ket step K steps (ket Hermite)
  static constexpr int NT = (NCD + 15) / 16;             // column tiles (cd)
  static constexpr int MTI = (NAB + 15) / 16;            // I row tiles (ab)
  static constexpr int KTI = 4 * MTG;                    // bra step K steps (bra Hermite, padded)
  static constexpr int NRL = (NR + 63) / 64;             // R entries per lane
  static constexpr int NEB = ne_axis(LA, LB), NEK = ne_axis(LC, LD);
  static constexpr int RSB = rec_size(LA, LB), RSK = rec_size(LC, LD);
};

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void eri_mfma_kernel(const int2* __restrict__ quartets, int first, int n, int splb,
                                                      const int4* __restrict__ pairs, const double* __restrict__ rec,
                                                      double* __restrict__ out, long long xs, long long qs,
                                                      long long ss, const double* __restrict__ gtab, double qcut) {
  using S = MfmaShape<LA, LB, LC, LD>;
  __shared__ double rbuf[2][S::NR + 1];
  const int lane = threadIdx.x;
  const int lr = lane & 15, lk = lane >> 4;
  // this lane's fixed shares: R-level entries, gather indices of the ket
  // A operand, packed E~ entries of the ket B operand and of the bra A operand
  int rent[S::NRL];
#pragma unroll
  for (int s = 0; s < S::NRL; ++s) {
    const int h = lane + 64 * s;
    int pk = 0;
    if (h > 0 && h < S::NR) {
      const int t = h_t(h), u = h_u(h), v = h_v(h);
      int axis, c, p1, p2;
      if (t > 0) {
        axis = 0, c = t - 1, p1 = hidx(t - 1, u, v), p2 = t > 1 ? hidx(t - 2, u, v) : S::NR;
      } else if (u > 0) {
        axis = 1, c = u - 1, p1 = hidx(t, u - 1, v), p2 = u > 1 ? hidx(t, u - 2, v) : S::NR;
      } else {
        axis = 2, c = v - 1, p1 = hidx(t, u, v - 1), p2 = v > 1 ? hidx(t, u, v - 2) : S::NR;
      }
      pk = p1 | p2 << 9 | c << 18 | axis << 22;
    }
    rent[s] = pk;
  }
  int ridx[S::MTG][S::KTG];
#pragma unroll
  for (int mt = 0; mt < S::MTG; ++mt)
#pragma unroll
    for (int ks = 0; ks < S::KTG; ++ks) {
      const int h = 16 * mt + lr, k = 4 * ks + lk;
      ridx[mt][ks] = (h < S::NHB && k < S::NHK)
                         ? hidx(h_t(h) + h_t(k), h_u(h) + h_u(k), h_v(h) + h_v(k))
                         : S::NR;
    }
  int eket[S::KTG][S::NT];
#pragma unroll
  for (int ks = 0; ks < S::KTG; ++ks)
#pragma unroll
    for (int nt = 0; nt < S::NT; ++nt)
      eket[ks][nt] = e_entry(LC, LD, 16 * nt + lr, 4 * ks + lk, S::NHK, S::NCD, true);
  int ebra[S::MTI][S::KTI];
#pragma unroll
  for (int mt = 0; mt < S::MTI; ++mt)
#pragma unroll
    for (int ks = 0; ks < S::KTI; ++ks)
      ebra[mt][ks] = e_entry(LA, LB, 16 * mt + lr, 4 * ks + lk, S::NHB, S::NAB, false);
  if (lane < 2) rbuf[lane][S::NR] = 0.0;
  __syncthreads();

  const int nw = n * splb;
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int sp = w / n;
    const int q = first + (w - sp * n);
    const int2 qd = quartets[q];
    const int4 pb = pairs[qd.x], pk = pairs[qd.y];
    const int szb = (pb.w + splb - 1) / splb;
    const int i0 = min(pb.w, sp * szb), i1 = min(pb.w, i0 + szb);
    const double* rb = rec + pb.z;
    const double* rk = rec + pk.z;
    const double sk0 = pk.w > 0 ? rk[5] : 0.0;
    mfma_d4 I[S::MTI][S::NT];
#pragma unroll
    for (int mt = 0; mt < S::MTI; ++mt)
#pragma unroll
      for (int nt = 0; nt < S::NT; ++nt) I[mt][nt] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    for (int i = i0; i < i1; ++i) {
      const double* bi = rb + i * S::RSB;
      const double sab = bi[5];
      // primitive-quartet cut (records in decreasing Schwarz factor)
      if (sab * sk0 < qcut) break;
      const double p = bi[0], Px = bi[1], Py = bi[2], Pz = bi[3], Kab = bi[4];
      mfma_d4 G[S::MTG][S::NT];
#pragma unroll
      for (int mt = 0; mt < S::MTG; ++mt)
#pragma unroll
        for (int nt = 0; nt < S::NT; ++nt) G[mt][nt] = mfma_d4{0.0, 0.0, 0.0, 0.0};
      for (int j = 0; j < pk.w; ++j) {
        const double* kj = rk + j * S::RSK;
        if (sab * kj[5] < qcut) break;
        const double qq = kj[0];
        const double X = Px - kj[1], Y = Py - kj[2], Z = Pz - kj[3];
        const double rs = rsqrt_nr(p + qq);
        const double alpha = p * qq * rs * rs;
        const double pref = TWO_PI_2P5 * Kab * kj[4] * rs;
        double F[S::L + 1];
        boys_global<S::L>(alpha * (X * X + Y * Y + Z * Z), gtab, F);
        // R^n_tuv from n = L down to 0, ping-pong in LDS
        const double m2a = -2.0 * alpha;
        double sn[S::L + 1];
        {
          double pw = 1.0;
#pragma unroll
          for (int m = 0; m <= S::L; ++m) {
            sn[m] = F[m] * pw;
            pw *= m2a;
          }
        }
        int cur = 0;
        if (lane == 0) rbuf[0][0] = sn[S::L];
        __syncthreads();
#pragma unroll
        for (int lev = S::L - 1; lev >= 0; --lev) {
          const int nxt = cur ^ 1;
          const int cnt = nherm(S::L - lev);
#pragma unroll
          for (int s = 0; s < S::NRL; ++s) {
            const int h = lane + 64 * s;
            if (h < cnt) {
              double v;
              if (h == 0) {
                v = sn[lev];
              } else {
                const int e = rent[s];
                const int p1 = e & 511, p2 = (e >> 9) & 511, c = (e >> 18) & 15, ax = e >> 22;
                const double d = ax == 0 ? X : (ax == 1 ? Y : Z);
                v = fma((double)c, rbuf[cur][p2], d * rbuf[cur][p1]);
              }
              rbuf[nxt][h] = v;
            }
          }
          __syncthreads();
          cur = nxt;
        }
        // ket step: G += R~ E~ (pref folded into E~)
#pragma unroll
        for (int ks = 0; ks < S::KTG; ++ks) {
          double b[S::NT];
#pragma unroll
          for (int nt = 0; nt < S::NT; ++nt) b[nt] = pref * eval_e(eket[ks][nt], kj + REC_HEAD, S::NEK);
#pragma unroll
          for (int mt = 0; mt < S::MTG; ++mt) {
            const double a = rbuf[cur][ridx[mt][ks]];
#pragma unroll
            for (int nt = 0; nt < S::NT; ++nt) G[mt][nt] = mfma16x16x4(a, b[nt], G[mt][nt]);
          }
        }
        __syncthreads();  // rbuf is rewritten by the next primitive quartet
      }
      // bra step: I += E~ G, G's result items are the B operands of K step 4 mt + r
#pragma unroll
      for (int mg = 0; mg < S::MTG; ++mg)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ks = 4 * mg + r;
#pragma unroll
          for (int mt = 0; mt < S::MTI; ++mt) {
            const double a = eval_e(ebra[mt][ks], bi + REC_HEAD, S::NEB);
#pragma unroll
            for (int nt = 0; nt < S::NT; ++nt) I[mt][nt] = mfma16x16x4(a, G[mg][nt][r], I[mt][nt]);
          }
        }
    }
    double* o = out + (long long)sp * ss + (long long)q * qs;
#pragma unroll
    for (int mt = 0; mt < S::MTI; ++mt)
#pragma unroll
      for (int nt = 0; nt < S::NT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ab = 16 * mt + 4 * r + lk, cd = 16 * nt + lr;
          if (ab < S::NAB && cd < S::NCD) o[(long long)(ab * S::NCD + cd) * xs] = I[mt][nt][r];
        }
  }
}

// the heavy combinations of the column kernel (kernels_chunk.hip)
#define DFT_FOR_MFMA(X) \
  X(8, 2, 0, 1, 1)      \
  X(11, 2, 1, 1, 0)     \
  X(12, 2, 1, 1, 1)     \
  X(13, 2, 1, 2, 0)     \
  X(14, 2, 1, 2, 1)     \
  X(16, 2, 2, 1, 0)     \
  X(17, 2, 2, 1, 1)     \
  X(18, 2, 2, 2, 0)     \
  X(19, 2, 2, 2, 1)     \
  X(20, 2, 2, 2, 2)

bool mfma_combo(int combo) {
  switch (combo) {
#define X(id, la, lb, lc, ld) \
  case id:                    \
    return true;
    DFT_FOR_MFMA(X)
#undef X
    default:
      return false;
  }
}

hipError_t launch_eri_mfma(int combo, const int2* q, int first, int n, int splb, const int4* pairs, const double* rec,
                           double* out, long long xs, long long qs, long long ss, const double* gtab, int cap,
                           bool cut, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long waves = (long long)n * splb;
  if (waves > 0x7fffffffLL) return hipErrorInvalidValue;
  const int grid = (int)std::min<long long>(waves, cap > 0 ? cap : 1 << 20);
  switch (combo) {
#define X(id, la, lb, lc, ld)                                                                                   \
  case id:                                                                                                    \
    hipLaunchKernelGGL((eri_mfma_kernel<la, lb, lc, ld>), dim3(grid), dim3(64), 0, s, q, first, n, splb, pairs, \
                       rec, out, xs, qs, ss, gtab, cut ? PRIM_QCUT64 : 0.0);                                  \
    return hipGetLastError();
    DFT_FOR_MFMA(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace dft
