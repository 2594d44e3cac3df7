// One-wave-per-quartet ERI (+ J/K digest) kernel for the heavy class
// combinations (d shells), where a lane-per-quartet kernel would need far
// more than the 512 VGPRs a lane can hold ((dd|dd) has 1296 integrals and
// 165 Hermite R values per primitive quartet).
//
// Per primitive quartet the 64 lanes of the wave
//   1. evaluate F_m (wave-uniform) and build R^n_tuv level by level in LDS,
//      each lane taking entries h = lane, lane+64, ... of a level;
//   2. ket step: lane-owned elements G[h][cd] (h in the bra Hermite set) are
//      accumulated over ket primitives in registers;
// after the ket loop G goes to LDS and the bra step accumulates the lane's
// integrals I[ab][cd]. After all primitives I goes to LDS and the lanes form
// the J/K contribution blocks (or, in raw mode, copy I out).
// Same arithmetic as the lane kernel (eri_quartet.h) with run-time class
// tables instead of compile-time unrolling.
#include <algorithm>

#include "kernels.h"
#include "md_device.h"

namespace dft {

constexpr int WMAX_L = 8;                 // (dd|dd)
constexpr int WMAX_NH = nherm(WMAX_L);    // 165
constexpr int WMAX_I = 6 * 6 * 6 * 6;     // 1296
constexpr int WMAX_G = nherm(4) * 36;     // 1260
constexpr int W_IPL = WMAX_I / 64 + 1;    // integrals per lane
constexpr int W_GPL = WMAX_G / 64 + 1;

struct WaveClass {
  int la, lb, lc, ld, na, nb, nc, nd, L, nhb, neb, nek, rsb, rsk, nG, nI;
  signed char comp[4][6][3];   // Cartesian exponents of the four shells
  short eob[6][6][3];          // E offset (per axis) of bra component pair (a, b)
  short eok[6][6][3];          // E offset of ket component pair (c, d)
};

__device__ __forceinline__ int hidx_rt(int t, int u, int v) { return hidx(t, u, v); }

template <typename real>
__device__ __forceinline__ real rec_e(const real* E, int off, int t) {
  // E_0^{00} = 1 by construction; off < 0 marks the (0,0) entry
  return off < 0 ? real(1) : E[off + t];
}

// Boys F_0..F_L from the global table (wave-uniform T), FP64 seed.
__device__ __forceinline__ void boys_global(int L, double T, const double* __restrict__ g, double* F) {
  if (T < BOYS_TMAX) {
    const int k = __double2int_rn(T * BOYS_INV_DT);
    const double d = -(T - k * BOYS_DT);
    const double inv_fact[BOYS_TAYLOR] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
    const double* c = g + k * BOYS_MT + L;
    double f = c[BOYS_TAYLOR - 1] * inv_fact[BOYS_TAYLOR - 1];
    for (int j = BOYS_TAYLOR - 2; j >= 0; --j) f = fma(f, d, c[j] * inv_fact[j]);
    F[L] = f;
    const double e = exp(-T), t2 = T + T;
    for (int m = L - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], e) * inv_odd(m);
  } else {
    F[0] = 0.88622692545275801365 * rsqrt(T);
    const double e = exp(-T), h = 0.5 / T;
    for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
  }
}

template <typename real, int NK, bool RAW>
__global__ __launch_bounds__(64) void eri_wave_kernel(WaveClass C, const int2* __restrict__ quartets,
                                                      const long long* __restrict__ slots, long long nq,
                                                      const int4* __restrict__ pairs, const real* __restrict__ rec,
                                                      const int4* __restrict__ shell_dm,
                                                      const double* __restrict__ DJ, const double* __restrict__ DK0,
                                                      const double* __restrict__ DK1,
                                                      double* __restrict__ out,
                                                      const double* __restrict__ boys_gtab) {
  __shared__ real sR[2][WMAX_NH];
  __shared__ real sGI[WMAX_I];
  __shared__ unsigned char sH[WMAX_NH][3];
  const int lane = threadIdx.x;
  for (int h = lane; h < WMAX_NH; h += 64) {
    sH[h][0] = (unsigned char)h_t(h);
    sH[h][1] = (unsigned char)h_u(h);
    sH[h][2] = (unsigned char)h_v(h);
  }
  __syncthreads();
  const int ncd = C.nc * C.nd, nab = C.na * C.nb;
  for (long long q = blockIdx.x; q < nq; q += gridDim.x) {
    const int2 qd = quartets[q];
    const int4 pb = pairs[qd.x];
    const int4 pk = pairs[qd.y];
    real Iacc[W_IPL];
#pragma unroll
    for (int k = 0; k < W_IPL; ++k) Iacc[k] = real(0);
    for (int i = 0; i < pb.w; ++i) {
      const real* rb = rec + pb.z + (long long)i * C.rsb;
      const real p = rb[0], Px = rb[1], Py = rb[2], Pz = rb[3], Kab = rb[4];
      real Gacc[W_GPL];
#pragma unroll
      for (int k = 0; k < W_GPL; ++k) Gacc[k] = real(0);
      for (int j = 0; j < pk.w; ++j) {
        const real* rk = rec + pk.z + (long long)j * C.rsk;
        const real qq = rk[0];
        const real X = Px - rk[1], Y = Py - rk[2], Z = Pz - rk[3];
        // K_ab and K_cd carry 1/p and 1/q (pair pre-pass)
        const real rs = rsqrt_nr(p + qq);
        const real alpha = p * qq * rs * rs;
        const real pref = real(34.986836655249725693) * Kab * rk[4] * rs;
        double F[WMAX_L + 2];
        boys_global(C.L, (double)(alpha * (X * X + Y * Y + Z * Z)), boys_gtab, F);
        // R levels: R^n_000 = (-2a)^n F_n
        const real m2a = real(-2) * alpha;
        real pw = real(1);
        for (int n = 0; n < C.L; ++n) pw *= m2a;
        if (lane == 0) sR[0][0] = (real)F[C.L] * pw;
        int cur = 0;
        __syncthreads();
        for (int n = C.L - 1; n >= 0; --n) {
          pw = real(1);
          for (int m = 0; m < n; ++m) pw *= m2a;
          const int S = C.L - n, nxt = cur ^ 1;
          for (int h = lane; h < nherm(S); h += 64) {
            const int t = sH[h][0], u = sH[h][1], v = sH[h][2];
            real r;
            if (h == 0) r = (real)F[n] * pw;
            else if (t > 0) {
              r = X * sR[cur][hidx_rt(t - 1, u, v)];
              if (t > 1) r += real(t - 1) * sR[cur][hidx_rt(t - 2, u, v)];
            } else if (u > 0) {
              r = Y * sR[cur][hidx_rt(t, u - 1, v)];
              if (u > 1) r += real(u - 1) * sR[cur][hidx_rt(t, u - 2, v)];
            } else {
              r = Z * sR[cur][hidx_rt(t, u, v - 1)];
              if (v > 1) r += real(v - 1) * sR[cur][hidx_rt(t, u, v - 2)];
            }
            sR[nxt][h] = r;
          }
          cur = nxt;
          __syncthreads();
        }
        const real* R = sR[cur];
        const real* Exk = rk + 5;
        const real* Eyk = Exk + C.nek;
        const real* Ezk = Eyk + C.nek;
#pragma unroll
        for (int k = 0; k < W_GPL; ++k) {
          const int g = lane + 64 * k;
          if (g < C.nG) {
            const int h = g / ncd, cd = g - h * ncd;
            const int c = cd / C.nd, d = cd - c * C.nd;
            const int t = sH[h][0], u = sH[h][1], v = sH[h][2];
            const int ex = C.comp[2][c][0] + C.comp[3][d][0], ey = C.comp[2][c][1] + C.comp[3][d][1],
                      ez = C.comp[2][c][2] + C.comp[3][d][2];
            const int ox = C.eok[c][d][0], oy = C.eok[c][d][1], oz = C.eok[c][d][2];
            real s = real(0);
            for (int ta = 0; ta <= ex; ++ta) {
              const real e1 = rec_e(Exk, ox, ta);
              for (int nu = 0; nu <= ey; ++nu) {
                const real e2 = e1 * rec_e(Eyk, oy, nu);
                for (int ph = 0; ph <= ez; ++ph) {
                  const real e3 = e2 * rec_e(Ezk, oz, ph);
                  const real r = R[hidx_rt(t + ta, u + nu, v + ph)];
                  s = ((ta + nu + ph) & 1) ? fma(-e3, r, s) : fma(e3, r, s);
                }
              }
            }
            Gacc[k] = fma(pref, s, Gacc[k]);
          }
        }
        __syncthreads();  // R buffers are rewritten by the next ket primitive
      }
      // bra step
#pragma unroll
      for (int k = 0; k < W_GPL; ++k) {
        const int g = lane + 64 * k;
        if (g < C.nG) sGI[g] = Gacc[k];
      }
      __syncthreads();
      const real* Exb = rb + 5;
      const real* Eyb = Exb + C.neb;
      const real* Ezb = Eyb + C.neb;
#pragma unroll
      for (int k = 0; k < W_IPL; ++k) {
        const int e = lane + 64 * k;
        if (e < C.nI) {
          const int ab = e / ncd, cd = e - ab * ncd;
          const int a = ab / C.nb, b = ab - a * C.nb;
          const int ex = C.comp[0][a][0] + C.comp[1][b][0], ey = C.comp[0][a][1] + C.comp[1][b][1],
                    ez = C.comp[0][a][2] + C.comp[1][b][2];
          const int ox = C.eob[a][b][0], oy = C.eob[a][b][1], oz = C.eob[a][b][2];
          real s = real(0);
          for (int t = 0; t <= ex; ++t) {
            const real e1 = rec_e(Exb, ox, t);
            for (int u = 0; u <= ey; ++u) {
              const real e2 = e1 * rec_e(Eyb, oy, u);
              for (int v = 0; v <= ez; ++v) s = fma(e2 * rec_e(Ezb, oz, v), sGI[hidx_rt(t, u, v) * ncd + cd], s);
            }
          }
          Iacc[k] += s;
        }
      }
      __syncthreads();
    }
    // integrals to LDS, then raw copy or J/K blocks
#pragma unroll
    for (int k = 0; k < W_IPL; ++k) {
      const int e = lane + 64 * k;
      if (e < C.nI) sGI[e] = Iacc[k];
    }
    __syncthreads();
    if constexpr (RAW) {
      for (int e = lane; e < C.nI; e += 64) out[q * C.nI + e] = (double)sGI[e];
    } else {
      double w = 1.0;
      if (pb.x == pb.y) w *= 0.5;
      if (pk.x == pk.y) w *= 0.5;
      if (qd.x == qd.y) w *= 0.5;
      const int4 sa = shell_dm[pb.x];
      const long long a0 = sa.x, b0 = shell_dm[pb.y].x, c0 = shell_dm[pk.x].x, d0 = shell_dm[pk.y].x;
      const long long nbf = sa.y;
      const long long mb = (long long)(unsigned)sa.z | ((long long)sa.w << 32);
      const double* DJm = DJ + mb;
      const long long* sl = slots + q * (2 + 4 * NK);
      const int na = C.na, nb = C.nb, nc = C.nc, nd = C.nd;
      auto I = [&](int a, int b, int c, int d) { return (double)sGI[((a * nb + b) * nc + c) * nd + d]; };
      const int nK = na * nc + na * nd + nb * nc + nb * nd;
      const int nout = nab + ncd + NK * nK;
      for (int o = lane; o < nout; o += 64) {
        double s = 0.0;
        if (o < nab) {
          const int a = o / nb, b = o - a * nb;
          for (int c = 0; c < nc; ++c)
            for (int d = 0; d < nd; ++d) s = fma(I(a, b, c, d), DJm[(c0 + c) * nbf + d0 + d], s);
          out[sl[0] + o] = 2.0 * w * s;
        } else if (o < nab + ncd) {
          const int x = o - nab, c = x / nd, d = x - c * nd;
          for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) s = fma(I(a, b, c, d), DJm[(a0 + a) * nbf + b0 + b], s);
          out[sl[1] + x] = 2.0 * w * s;
        } else {
          int x = o - nab - ncd;
          const int kk = x / nK;
          x -= kk * nK;
          const double* D = (kk == 0 ? DK0 : DK1) + mb;
          const long long* ks = sl + 2 + 4 * kk;
          if (x < na * nc) {
            const int a = x / nc, c = x - a * nc;
            for (int b = 0; b < nb; ++b)
              for (int d = 0; d < nd; ++d) s = fma(I(a, b, c, d), D[(b0 + b) * nbf + d0 + d], s);
            out[ks[0] + x] = w * s;
          } else if ((x -= na * nc) < na * nd) {
            const int a = x / nd, d = x - a * nd;
            for (int b = 0; b < nb; ++b)
              for (int c = 0; c < nc; ++c) s = fma(I(a, b, c, d), D[(b0 + b) * nbf + c0 + c], s);
            out[ks[1] + x] = w * s;
          } else if ((x -= na * nd) < nb * nc) {
            const int b = x / nc, c = x - b * nc;
            for (int a = 0; a < na; ++a)
              for (int d = 0; d < nd; ++d) s = fma(I(a, b, c, d), D[(a0 + a) * nbf + d0 + d], s);
            out[ks[2] + x] = w * s;
          } else {
            x -= nb * nc;
            const int b = x / nd, d = x - b * nd;
            for (int a = 0; a < na; ++a)
              for (int c = 0; c < nc; ++c) s = fma(I(a, b, c, d), D[(a0 + a) * nbf + c0 + c], s);
            out[ks[3] + x] = w * s;
          }
        }
      }
    }
    __syncthreads();  // sGI is reused by the next quartet
  }
}

static WaveClass make_class(int la, int lb, int lc, int ld) {
  WaveClass C{};
  const int ls[4] = {la, lb, lc, ld};
  C.la = la; C.lb = lb; C.lc = lc; C.ld = ld;
  C.na = ncart(la); C.nb = ncart(lb); C.nc = ncart(lc); C.nd = ncart(ld);
  C.L = la + lb + lc + ld;
  C.nhb = nherm(la + lb);
  C.neb = ne_axis(la, lb);
  C.nek = ne_axis(lc, ld);
  C.rsb = rec_size(la, lb);
  C.rsk = rec_size(lc, ld);
  C.nG = C.nhb * C.nc * C.nd;
  C.nI = C.na * C.nb * C.nc * C.nd;
  for (int s = 0; s < 4; ++s)
    for (int k = 0; k < ncart(ls[s]); ++k)
      for (int d = 0; d < 3; ++d) C.comp[s][k][d] = (signed char)comp_of(ls[s], k, d);
  for (int a = 0; a < C.na; ++a)
    for (int b = 0; b < C.nb; ++b)
      for (int d = 0; d < 3; ++d) {
        const int ia = C.comp[0][a][d], ib = C.comp[1][b][d];
        C.eob[a][b][d] = (short)((ia == 0 && ib == 0) ? -1 : eoff(lb, ia, ib));
      }
  for (int c = 0; c < C.nc; ++c)
    for (int e = 0; e < C.nd; ++e)
      for (int d = 0; d < 3; ++d) {
        const int ic = C.comp[2][c][d], id = C.comp[3][e][d];
        C.eok[c][e][d] = (short)((ic == 0 && id == 0) ? -1 : eoff(ld, ic, id));
      }
  return C;
}

hipError_t launch_eri_wave(int la, int lb, int lc, int ld, bool fp32, bool raw, const int2* q, const long long* sl,
                           long long nq, const int4* pairs, const void* rec, const int4* ao, const double* DJ,
                           const double* DK0, const double* DK1, int nk, double* out, const double* bt,
                           int cap, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  if (la + lb + lc + ld > WMAX_L || la > 2 || lb > 2 || lc > 2 || ld > 2) return hipErrorInvalidValue;
  const WaveClass C = make_class(la, lb, lc, ld);
  dim3 grid((unsigned)std::min<long long>(nq, cap)), block(64);
  if (raw) {
    hipLaunchKernelGGL((eri_wave_kernel<double, 1, true>), grid, block, 0, s, C, q, sl, nq, pairs,
                       (const double*)rec, ao, DJ, DK0, DK1, out, bt);
  } else if (fp32) {
    if (nk == 1)
      hipLaunchKernelGGL((eri_wave_kernel<float, 1, false>), grid, block, 0, s, C, q, sl, nq, pairs,
                         (const float*)rec, ao, DJ, DK0, DK1, out, bt);
    else
      hipLaunchKernelGGL((eri_wave_kernel<float, 2, false>), grid, block, 0, s, C, q, sl, nq, pairs,
                         (const float*)rec, ao, DJ, DK0, DK1, out, bt);
  } else {
    if (nk == 1)
      hipLaunchKernelGGL((eri_wave_kernel<double, 1, false>), grid, block, 0, s, C, q, sl, nq, pairs,
                         (const double*)rec, ao, DJ, DK0, DK1, out, bt);
    else
      hipLaunchKernelGGL((eri_wave_kernel<double, 2, false>), grid, block, 0, s, C, q, sl, nq, pairs,
                         (const double*)rec, ao, DJ, DK0, DK1, out, bt);
  }
  return hipGetLastError();
}

}  // namespace dft
