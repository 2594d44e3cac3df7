// Shell quartets of any class up to (ff|ff), with the momenta as run-time
// values: the integrals of the class combinations that contain an f shell
// (no compile-time kernel is instantiated for them) and their J/K blocks.
//
// One 64-lane workgroup (one wave) per contracted quartet. Per primitive
// quartet the wave builds R^0_tuv (t+u+v <= L <= 12) level by level in LDS
// (entries of a level in parallel over the lanes), then
//   ket step  G[h][cd] += pref sum_{tau nu phi} (-1)^{tau+nu+phi}
//                          E^cd_tau E^cd_nu E^cd_phi R[h + (tau,nu,phi)]
// over (h, cd) in parallel, and once per bra primitive the bra step
//   I[ab][cd] += sum_h E^ab_t E^ab_u E^ab_v G[h][cd]
// over (ab, cd) in parallel; G and I live in LDS (ni <= 10^4 doubles). The
// arithmetic is that of eri_quartet.h (McMurchie-Davidson with shell-level
// shared R; pref = 2 pi^{5/2} K_ab K_cd / sqrt(p+q)); the reference computes
// the same integrals per primitive Cartesian quartet (eri_kernel,
// shelled.py:452-532) for momenta up to its L = 12 capacity
// (shelled.py:168-171).
#include <algorithm>

#include "eri_quartet.h"
#include "kernels.h"

namespace dft {

namespace {

constexpr int GEN_LMAX = 3;
constexpr int GEN_L = 4 * GEN_LMAX;
constexpr int GEN_NH = nherm(GEN_L);  // 455
constexpr int GEN_NE = ne_axis(GEN_LMAX, GEN_LMAX);
constexpr int GEN_NC = ncart(GEN_LMAX);
constexpr int GEN_LDS_MAX = 160 * 1024;

// (t, u, v) of Hermite index h, packed t | u << 8 | v << 16
struct HermTable {
  int tuv[GEN_NH];
  constexpr HermTable() : tuv() {
    for (int h = 0; h < GEN_NH; ++h) tuv[h] = h_t(h) | h_u(h) << 8 | h_v(h) << 16;
  }
};
__constant__ HermTable c_herm = HermTable();

// Cartesian component powers of l <= GEN_LMAX: packed x | y << 8 | z << 16
struct CompTable {
  int xyz[GEN_LMAX + 1][GEN_NC];
  constexpr CompTable() : xyz() {
    for (int l = 0; l <= GEN_LMAX; ++l)
      for (int k = 0; k < ncart(l); ++k) xyz[l][k] = comp_x(l, k) | comp_y(l, k) << 8 | comp_z(l, k) << 16;
  }
};
__constant__ CompTable c_comp = CompTable();

__device__ __forceinline__ int px(int w) { return w & 255; }
__device__ __forceinline__ int py(int w) { return (w >> 8) & 255; }
__device__ __forceinline__ int pz(int w) { return w >> 16; }

// offset of E_0^{ia,ib} in a record's axis table (eoff of md_device.h)
__device__ __forceinline__ int eoff_rt(int lb, int ia, int ib) {
  int o = 0;
  for (int a = 0; a < ia; ++a) o += (lb + 1) * (a + 1) + lb * (lb + 1) / 2;
  return o + ib * (ia + 1) + ib * (ib - 1) / 2;
}

struct GenShape {
  int la, lb, lc, ld, L, na, nb, nc, nd, nab, ncd, ni, nhb, neb, nek, rsb, rsk;
  __host__ __device__ GenShape(int a, int b, int c, int d)
      : la(a), lb(b), lc(c), ld(d), L(a + b + c + d), na(ncart(a)), nb(ncart(b)), nc(ncart(c)), nd(ncart(d)),
        nab(ncart(a) * ncart(b)), ncd(ncart(c) * ncart(d)), ni(ncart(a) * ncart(b) * ncart(c) * ncart(d)),
        nhb(nherm(a + b)), neb(ne_axis(a, b)), nek(ne_axis(c, d)), rsb(rec_size(a, b)), rsk(rec_size(c, d)) {}
  // LDS doubles: bra and ket E tables, F, two R levels, G, I
  __host__ __device__ int lds_doubles() const { return 3 * neb + 3 * nek + 16 + 2 * nherm(L) + nhb * ncd + ni; }
};

}  // namespace

// Integrals of quartets [first, first + n) of a list (plain FP64 records:
// pair (sh_a, sh_b, record offset, primitive pairs)), written to
// out[x * xs + q * qs] (x = abcd row-major), q the index in the list.
__global__ __launch_bounds__(64) void eri_generic_kernel(int la, int lb, int lc, int ld,
                                                         const int2* __restrict__ quartets, int first, int n,
                                                         const int4* __restrict__ pairs,
                                                         const double* __restrict__ rec, double* __restrict__ out,
                                                         long long xs, long long qs,
                                                         const double* __restrict__ boys_gtab, bool cut) {
  extern __shared__ double lds[];
  const GenShape S(la, lb, lc, ld);
  const int lane = threadIdx.x;
  double* Eb = lds;                       // 3 * neb
  double* Ek = Eb + 3 * S.neb;            // 3 * nek
  double* F = Ek + 3 * S.nek;             // L + 1 <= 13 (16 reserved)
  double* R0 = F + 16;                    // two levels of nherm(L)
  double* R1 = R0 + nherm(S.L);
  double* G = R1 + nherm(S.L);            // nhb * ncd
  double* I = G + S.nhb * S.ncd;          // ni
  const int nhL = nherm(S.L);
  for (int qi = blockIdx.x; qi < n; qi += gridDim.x) {
    const int q = first + qi;
    const int2 qd = quartets[q];
    const int4 pb = pairs[qd.x], pk = pairs[qd.y];
    for (int x = lane; x < S.ni; x += 64) I[x] = 0.0;
    const double sk0 = cut && pk.w > 0 ? rec[pk.z + 5] : 0.0;  // largest ket Schwarz factor
    for (int i = 0; i < pb.w; ++i) {
      const double* rb = rec + pb.z + (long long)i * S.rsb;
      const double p = rb[0], Px = rb[1], Py = rb[2], Pz = rb[3], Kab = rb[4], sab = rb[5];
      if (cut && sab * sk0 < PRIM_QCUT64) break;
      for (int x = lane; x < 3 * S.neb; x += 64) Eb[x] = rb[REC_HEAD + x];
      for (int x = lane; x < S.nhb * S.ncd; x += 64) G[x] = 0.0;
      for (int j = 0; j < pk.w; ++j) {
        const double* rk = rec + pk.z + (long long)j * S.rsk;
        const double qq = rk[0], Qx = rk[1], Qy = rk[2], Qz = rk[3], Kcd = rk[4], scd = rk[5];
        if (cut && sab * scd < PRIM_QCUT64) break;  // ket primitives are in decreasing s
        __syncthreads();  // the previous primitive quartet's readers of Ek, R are done
        for (int x = lane; x < 3 * S.nek; x += 64) Ek[x] = rk[REC_HEAD + x];
        const double rs = 1.0 / sqrt(p + qq);
        const double alpha = p * qq * rs * rs;
        const double pref = TWO_PI_2P5 * Kab * Kcd * rs;
        const double X = Px - Qx, Y = Py - Qy, Z = Pz - Qz;
        const double T = alpha * (X * X + Y * Y + Z * Z);
        if (lane == 0) {
          // F_0..F_L: the Taylor step of the tabulated F_L, downward
          // recursion; the asymptotic form above the table
          if (T < BOYS_TMAX) {
            const int k = boys_point(T);
            const double dd = -(T - k * BOYS_DT);
            const double inv_fact[BOYS_TAYLOR] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
            const double* row = boys_gtab + k * BOYS_MT + S.L;
            double f = row[BOYS_TAYLOR - 1] * inv_fact[BOYS_TAYLOR - 1];
            for (int jj = BOYS_TAYLOR - 2; jj >= 0; --jj) f = fma(f, dd, row[jj] * inv_fact[jj]);
            F[S.L] = f;
            const double e = exp(-T), t2 = T + T;
            for (int m = S.L - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], e) * inv_odd(m);
          } else {
            const double r = 1.0 / sqrt(T), h = 0.5 * r * r;
            const double e = T >= boys_tnoexp(S.L) ? 0.0 : exp(-T);
            F[0] = 0.88622692545275801365 * r;
            for (int m = 0; m < S.L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
          }
          // R^n_000 = (-2 alpha)^n F_n
          double pw = 1.0;
          for (int m = 0; m <= S.L; ++m) {
            F[m] *= pw;
            pw *= -2.0 * alpha;
          }
          R0[0] = F[S.L];
        }
        __syncthreads();
        // levels n = L-1 .. 0: R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{tuv}
        double* cur = R0;
        double* nxt = R1;
        for (int lev = 1; lev <= S.L; ++lev) {
          const int nh = nherm(lev);
          for (int h = lane; h < nh; h += 64) {
            double r;
            if (h == 0) {
              r = F[S.L - lev];
            } else {
              const int w = c_herm.tuv[h], t = px(w), u = py(w), v = pz(w);
              if (t > 0) {
                r = X * cur[hidx(t - 1, u, v)];
                if (t > 1) r = fma(double(t - 1), cur[hidx(t - 2, u, v)], r);
              } else if (u > 0) {
                r = Y * cur[hidx(t, u - 1, v)];
                if (u > 1) r = fma(double(u - 1), cur[hidx(t, u - 2, v)], r);
              } else {
                r = Z * cur[hidx(t, u, v - 1)];
                if (v > 1) r = fma(double(v - 1), cur[hidx(t, u, v - 2)], r);
              }
            }
            nxt[h] = r;
          }
          __syncthreads();
          double* tmp = cur;
          cur = nxt;
          nxt = tmp;
        }
        // cur: R^0_tuv, t+u+v <= L
        for (int x = lane; x < S.nhb * S.ncd; x += 64) {
          const int h = x / S.ncd, cdi = x - h * S.ncd;
          const int c = cdi / S.nd, d = cdi - c * S.nd;
          const int wh = c_herm.tuv[h], wc = c_comp.xyz[lc][c], wd = c_comp.xyz[ld][d];
          const int t = px(wh), u = py(wh), v = pz(wh);
          const int cx = px(wc), cy = py(wc), cz = pz(wc), dx = px(wd), dy = py(wd), dz = pz(wd);
          const double* ex = Ek + eoff_rt(ld, cx, dx);
          const double* ey = Ek + S.nek + eoff_rt(ld, cy, dy);
          const double* ez = Ek + 2 * S.nek + eoff_rt(ld, cz, dz);
          double s = 0.0;
          for (int ta = 0; ta <= cx + dx; ++ta) {
            const double exv = (ta == 0 && cx + dx == 0) ? 1.0 : ex[ta];
            for (int nu = 0; nu <= cy + dy; ++nu) {
              const double eyv = (nu == 0 && cy + dy == 0) ? 1.0 : ey[nu];
              double s2 = 0.0;
              for (int ph = 0; ph <= cz + dz; ++ph) {
                const double ezv = (ph == 0 && cz + dz == 0) ? 1.0 : ez[ph];
                const double r = cur[hidx(t + ta, u + nu, v + ph)];
                s2 = (ph & 1) ? fma(-ezv, r, s2) : fma(ezv, r, s2);
              }
              s = ((ta + nu) & 1) ? fma(-exv * eyv, s2, s) : fma(exv * eyv, s2, s);
            }
          }
          G[x] = fma(pref, s, G[x]);
        }
      }
      __syncthreads();  // G complete
      for (int x = lane; x < S.ni; x += 64) {
        const int ab = x / S.ncd, cdi = x - ab * S.ncd;
        const int a = ab / S.nb, b = ab - a * S.nb;
        const int wa = c_comp.xyz[la][a], wb = c_comp.xyz[lb][b];
        const int ax = px(wa), ay = py(wa), az = pz(wa), bx = px(wb), by = py(wb), bz = pz(wb);
        const double* ex = Eb + eoff_rt(lb, ax, bx);
        const double* ey = Eb + S.neb + eoff_rt(lb, ay, by);
        const double* ez = Eb + 2 * S.neb + eoff_rt(lb, az, bz);
        double s = 0.0;
        for (int t = 0; t <= ax + bx; ++t) {
          const double exv = (t == 0 && ax + bx == 0) ? 1.0 : ex[t];
          for (int u = 0; u <= ay + by; ++u) {
            const double eyv = (u == 0 && ay + by == 0) ? 1.0 : ey[u];
            double s2 = 0.0;
            for (int v = 0; v <= az + bz; ++v) {
              const double ezv = (v == 0 && az + bz == 0) ? 1.0 : ez[v];
              s2 = fma(ezv, G[hidx(t, u, v) * S.ncd + cdi], s2);
            }
            s = fma(exv * eyv, s2, s);
          }
        }
        I[x] += s;
      }
      __syncthreads();  // the bra step's reads of G and Eb are done
    }
    for (int x = lane; x < S.ni; x += 64) out[x * xs + (long long)q * qs] = I[x];
    __syncthreads();
  }
  (void)nhL;
}

// J/K blocks of quartets [0, n) from their integrals (ibuf + q * ni), one wave
// per quartet, the same blocks and weights as the lane kernels (Jab, Jcd,
// then per exchange density Kac, Kad, Kbc, Kbd; w = 2^-([a=b]+[c=d]+[ab=cd]))
// written to stage at the quartet's slots (quartet-major, 2 + 4 nk per
// quartet). Integrals and densities are read from global memory.
template <int NK>
__global__ __launch_bounds__(256) void jk_generic_kernel(int na, int nb, int nc, int nd,
                                                         const int2* __restrict__ quartets,
                                                         const long long* __restrict__ slots, int n,
                                                         const int4* __restrict__ pairs,
                                                         const int4* __restrict__ shell_dm,
                                                         const double* __restrict__ DJ,
                                                         const double* __restrict__ DK0,
                                                         const double* __restrict__ DK1,
                                                         const double* __restrict__ ibuf, double* __restrict__ stage) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int nab = na * nb, ncd = nc * nd, ni = nab * ncd;
  const int nK = na * nc + na * nd + nb * nc + nb * nd;
  const int nout = nab + ncd + NK * nK;
  const int wv = threadIdx.x >> 6;
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
    const long long* sl = slots + (long long)q * (2 + 4 * NK);
    for (int o = lane; o < nout; o += 64) {
      double s = 0.0;
      int blk, x;
      double scale;
      if (o < nab) {  // J_ab = 2w sum_cd I P_cd
        const int a = o / nb, b = o - a * nb;
        const double* D = DJ + mb;
        for (int c = 0; c < nc; ++c)
          for (int d = 0; d < nd; ++d) s = fma(Iq[(a * nb + b) * ncd + c * nd + d], D[(c0 + c) * nbf + d0 + d], s);
        blk = 0, x = o, scale = 2.0 * w;
      } else if (o < nab + ncd) {  // J_cd = 2w sum_ab I P_ab
        x = o - nab;
        const int c = x / nd, d = x - c * nd;
        const double* D = DJ + mb;
        for (int a = 0; a < na; ++a)
          for (int b = 0; b < nb; ++b) s = fma(Iq[(a * nb + b) * ncd + c * nd + d], D[(a0 + a) * nbf + b0 + b], s);
        blk = 1, scale = 2.0 * w;
      } else {
        x = o - nab - ncd;
        const int kk = x / nK;
        x -= kk * nK;
        const double* D = (kk == 0 ? DK0 : DK1) + mb;
        const int t0 = 2 + 4 * kk;
        scale = w;
        if (x < na * nc) {  // K_ac = w sum_bd I P_bd
          const int a = x / nc, c = x - a * nc;
          for (int b = 0; b < nb; ++b)
            for (int d = 0; d < nd; ++d) s = fma(Iq[(a * nb + b) * ncd + c * nd + d], D[(b0 + b) * nbf + d0 + d], s);
          blk = t0;
        } else if ((x -= na * nc) < na * nd) {  // K_ad = w sum_bc I P_bc
          const int a = x / nd, d = x - a * nd;
          for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c) s = fma(Iq[(a * nb + b) * ncd + c * nd + d], D[(b0 + b) * nbf + c0 + c], s);
          blk = t0 + 1;
        } else if ((x -= na * nd) < nb * nc) {  // K_bc = w sum_ad I P_ad
          const int b = x / nc, c = x - b * nc;
          for (int a = 0; a < na; ++a)
            for (int d = 0; d < nd; ++d) s = fma(Iq[(a * nb + b) * ncd + c * nd + d], D[(a0 + a) * nbf + d0 + d], s);
          blk = t0 + 2;
        } else {  // K_bd = w sum_ac I P_ac
          x -= nb * nc;
          const int b = x / nd, d = x - b * nd;
          for (int a = 0; a < na; ++a)
            for (int c = 0; c < nc; ++c) s = fma(Iq[(a * nb + b) * ncd + c * nd + d], D[(a0 + a) * nbf + c0 + c], s);
          blk = t0 + 3;
        }
      }
      stage[sl[blk] + x] = scale * s;
    }
  }
}

bool generic_combo(int combo) {
  int la, lb, lc, ld;
  pc_moments(combo_pcb(combo), la, lb);
  pc_moments(combo_pck(combo), lc, ld);
  return std::max(std::max(la, lb), std::max(lc, ld)) == GEN_LMAX;
}

size_t generic_lds_bytes(int la, int lb, int lc, int ld) {
  return (size_t)GenShape(la, lb, lc, ld).lds_doubles() * sizeof(double);
}

hipError_t launch_eri_generic(int la, int lb, int lc, int ld, const int2* q, int first, int n, const int4* pairs,
                              const double* rec, double* out, long long xs, long long qs, const double* gtab,
                              int cap, bool cut, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (std::max(std::max(la, lb), std::max(lc, ld)) > GEN_LMAX || std::min(std::min(la, lb), std::min(lc, ld)) < 0)
    return hipErrorInvalidValue;
  const size_t lds = generic_lds_bytes(la, lb, lc, ld);
  if (lds > (size_t)GEN_LDS_MAX) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)eri_generic_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       GEN_LDS_MAX);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const int grid = std::max(1, std::min(n, cap > 0 ? cap : 65536));
  hipLaunchKernelGGL(eri_generic_kernel, dim3(grid), dim3(64), lds, s, la, lb, lc, ld, q, first, n, pairs, rec, out,
                     xs, qs, gtab, cut);
  return hipGetLastError();
}

hipError_t launch_jk_generic(int la, int lb, int lc, int ld, const int2* q, const long long* slots, int n,
                             const int4* pairs, const int4* shell_dm, const double* DJ, const double* DK0,
                             const double* DK1, int nk, const double* ibuf, double* stage, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int wpb = 4;
  dim3 grid((unsigned)std::min<long long>(((long long)n + wpb - 1) / wpb, 65536)), blk(64 * wpb);
  if (nk == 1)
    hipLaunchKernelGGL((jk_generic_kernel<1>), grid, blk, 0, s, ncart(la), ncart(lb), ncart(lc), ncart(ld), q, slots,
                       n, pairs, shell_dm, DJ, DK0, DK1, ibuf, stage);
  else
    hipLaunchKernelGGL((jk_generic_kernel<2>), grid, blk, 0, s, ncart(la), ncart(lb), ncart(lc), ncart(ld), q, slots,
                       n, pairs, shell_dm, DJ, DK0, DK1, ibuf, stage);
  return hipGetLastError();
}

}  // namespace dft
