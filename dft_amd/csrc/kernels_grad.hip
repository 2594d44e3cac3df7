// Analytic nuclear-gradient kernels (s/p shells).
//
// Two-electron part. For a canonical quartet with degeneracy weight w the
// energy is sum_{abcd} Gamma_abcd (ab|cd) with the density pair
//   Gamma_abcd = w [4 P_ab P_cd - 2 sum_s (P^s_ac P^s_bd + P^s_ad P^s_bc)],
// so dE/dA_x = sum Gamma_abcd d(ab|cd)/dA_x. The derivative of a Cartesian
// Gaussian, d/dA_x x_A^i e^{-a x_A^2} = 2a x_A^{i+1} e - i x_A^{i-1} e, gives
// the Hermite E derivative recurrence
//   dE^{ij}_t / dA_x = 2a E^{i+1,j}_t - i E^{i-1,j}_t   (dB: 2b E^{i,j+1}_t - j E^{i,j-1}_t)
// used here with E tables extended by one in each index. One launch gives
// the derivatives with respect to the two bra centres; the same quartets
// with bra and ket exchanged give the ket centres (Gamma is symmetric under
// the exchange).
//
// One-electron part: per (shell pair, nucleus C) block, the derivatives of
// S (weighted by W), T (by P) and of the nuclear attraction of C (by P) with
// respect to A and B, and the operator derivative with respect to C
// (dV^C/dC = -(dV^C/dA + dV^C/dB), translational invariance).
//
// Every contribution is written to its own staging slot; a fixed-order
// per-atom reduction sums them (no atomics).
#include "eri_quartet.h"
#include "kernels.h"

namespace dft {

// gradient pair record: p, Px, Py, Pz, Kab, alpha, then E tables of class
// (la+1, lb+1) without the angular normalisation (s/p functions have none)
__host__ __device__ constexpr int grec_size(int la, int lb) { return 6 + 3 * ne_axis(la + 1, lb + 1); }

template <int LB1, int I, int J, int T>
__device__ __forceinline__ double eg(const double* __restrict__ E) {
  if constexpr (I < 0 || J < 0 || T < 0 || T > I + J) return 0.0;
  else if constexpr (I == 0 && J == 0) return 1.0;
  else return E[eoff(LB1, I, J) + T];
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void eri_grad_kernel(const int2* __restrict__ quartets,
                                                       const long long* __restrict__ slots, long long nq,
                                                       const int4* __restrict__ gpairs, const double* __restrict__ grec,
                                                       const int* __restrict__ shell_ao, const double* __restrict__ DJ,
                                                       const double* __restrict__ DK0, const double* __restrict__ DK1,
                                                       int nbf, double* __restrict__ stage,
                                                       const double* __restrict__ boys_gtab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  constexpr int L1 = LA + LB + LC + LD + 1;
  constexpr int NH1 = nherm(LA + LB + 1);
  constexpr int RSB = grec_size(LA, LB), RSK = grec_size(LC, LD);
  constexpr int NEB = ne_axis(LA + 1, LB + 1), NEK = ne_axis(LC + 1, LD + 1);
  __shared__ double btab[BOYS_LDS];
  boys_stage<L1>(boys_gtab, btab);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += stride) {
    const int2 qd = quartets[q];
    const int4 pb = gpairs[qd.x];
    const int4 pk = gpairs[qd.y];
    // the canonical weight is symmetric in bra/ket, so it is the same in both launches
    double w = 1.0;
    if (pb.x == pb.y) w *= 0.5;
    if (pk.x == pk.y) w *= 0.5;
    if (qd.x == qd.y) w *= 0.5;
    const long long a0 = shell_ao[pb.x], b0 = shell_ao[pb.y], c0 = shell_ao[pk.x], d0 = shell_ao[pk.y];
    const double* DB = DK1 ? DK1 : DK0;
    double Gam[NAB * NCD];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            const long long A = a0 + a, B = b0 + b, Cc = c0 + c, D = d0 + d;
            const double x = DK0[A * nbf + Cc] * DK0[B * nbf + D] + DK0[A * nbf + D] * DK0[B * nbf + Cc] +
                             DB[A * nbf + Cc] * DB[B * nbf + D] + DB[A * nbf + D] * DB[B * nbf + Cc];
            Gam[((a * NB + b) * NC + c) * ND + d] = w * (4.0 * DJ[A * nbf + B] * DJ[Cc * nbf + D] - 2.0 * x);
          }
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < pb.w; ++i) {
      const double* rb = grec + pb.z + (long long)i * RSB;
      const double p = rb[0], Px = rb[1], Py = rb[2], Pz = rb[3], Kab = rb[4], al = rb[5], be = p - al;
      double G[NH1 * NCD];
#pragma unroll
      for (int x = 0; x < NH1 * NCD; ++x) G[x] = 0.0;
      for (int j = 0; j < pk.w; ++j) {
        const double* rk = grec + pk.z + (long long)j * RSK;
        const double qq = rk[0];
        const double X = Px - rk[1], Y = Py - rk[2], Z = Pz - rk[3];
        const double pq = p + qq;
        const double alpha = p * qq / pq;
        const double pref = TWO_PI_2P5 * Kab * rk[4] / (p * qq) * sqrt(1.0 / pq);
        double F[L1 + 1];
        boys_fp64<L1>(alpha * (X * X + Y * Y + Z * Z), btab, F);
        double R[nherm(L1)];
        hermite_R<L1, double>(F, -2.0 * alpha, X, Y, Z, R);
        const double* Exk = rk + 6;
        const double* Eyk = Exk + NEK;
        const double* Ezk = Eyk + NEK;
        static_for<0, NCD>([&](auto cdI) {
          constexpr int cd = decltype(cdI)::value;
          constexpr int c = cd / ND, d = cd % ND;
          constexpr int cx = comp_x(LC, c), cy = comp_y(LC, c), cz = comp_z(LC, c);
          constexpr int dx = comp_x(LD, d), dy = comp_y(LD, d), dz = comp_z(LD, d);
          static_for<0, NH1>([&](auto hI) {
            constexpr int h = decltype(hI)::value;
            constexpr int t = h_t(h), u = h_u(h), v = h_v(h);
            double s = 0.0;
            static_for<0, cx + dx + 1>([&](auto tI) {
              static_for<0, cy + dy + 1>([&](auto uI) {
                static_for<0, cz + dz + 1>([&](auto vI) {
                  constexpr int ta = decltype(tI)::value, nu = decltype(uI)::value, ph = decltype(vI)::value;
                  const double e = eg<LD + 1, cx, dx, ta>(Exk) * eg<LD + 1, cy, dy, nu>(Eyk) *
                                   eg<LD + 1, cz, dz, ph>(Ezk);
                  const double r = R[hidx(t + ta, u + nu, v + ph)];
                  s = ((ta + nu + ph) & 1) ? fma(-e, r, s) : fma(e, r, s);
                });
              });
            });
            G[h * NCD + cd] = fma(pref, s, G[h * NCD + cd]);
          });
        });
      }
      const double* Exb = rb + 6;
      const double* Eyb = Exb + NEB;
      const double* Ezb = Eyb + NEB;
      static_for<0, NAB>([&](auto abI) {
        constexpr int ab = decltype(abI)::value;
        constexpr int a = ab / NB, b = ab % NB;
        constexpr int ai[3] = {comp_x(LA, a), comp_y(LA, a), comp_z(LA, a)};
        constexpr int bi[3] = {comp_x(LB, b), comp_y(LB, b), comp_z(LB, b)};
        // Gt[h] = sum_cd Gamma_abcd G[h][cd]
        double Gt[NH1];
#pragma unroll
        for (int h = 0; h < NH1; ++h) {
          double s = 0.0;
#pragma unroll
          for (int cd = 0; cd < NCD; ++cd) s = fma(Gam[ab * NCD + cd], G[h * NCD + cd], s);
          Gt[h] = s;
        }
        const double* Eax[3] = {Exb, Eyb, Ezb};
        static_for<0, 3>([&](auto axI) {
          constexpr int ax = decltype(axI)::value;
          constexpr int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
          // undifferentiated E along the other two axes times the derivative along ax
          static_for<0, ai[ax] + bi[ax] + 2>([&](auto tI) {
            constexpr int t = decltype(tI)::value;
            const double dA = 2.0 * al * eg<LB + 1, ai[ax] + 1, bi[ax], t>(Eax[ax]) -
                              (ai[ax] > 0 ? ai[ax] * eg<LB + 1, ai[ax] - 1, bi[ax], t>(Eax[ax]) : 0.0);
            const double dB = 2.0 * be * eg<LB + 1, ai[ax], bi[ax] + 1, t>(Eax[ax]) -
                              (bi[ax] > 0 ? bi[ax] * eg<LB + 1, ai[ax], bi[ax] - 1, t>(Eax[ax]) : 0.0);
            static_for<0, ai[o1] + bi[o1] + 1>([&](auto uI) {
              static_for<0, ai[o2] + bi[o2] + 1>([&](auto vI) {
                constexpr int u = decltype(uI)::value, v = decltype(vI)::value;
                const double e = eg<LB + 1, ai[o1], bi[o1], u>(Eax[o1]) * eg<LB + 1, ai[o2], bi[o2], v>(Eax[o2]);
                // (t,u,v) along (ax, o1, o2) -> (x, y, z) indices
                constexpr int hx = ax == 0 ? t : (o1 == 0 ? u : v);
                constexpr int hy = ax == 1 ? t : (o1 == 1 ? u : v);
                constexpr int hz = ax == 2 ? t : (o1 == 2 ? u : v);
                const double g = Gt[hidx(hx, hy, hz)] * e;
                gA[ax] = fma(dA, g, gA[ax]);
                gB[ax] = fma(dB, g, gB[ax]);
              });
            });
          });
        });
      });
    }
    const long long* sl = slots + 2 * q;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      stage[sl[0] + x] = gA[x];
      stage[sl[1] + x] = gB[x];
    }
  }
}

// gradient pair records: one lane per primitive pair, E up to (la+1, lb+1)
__global__ void grad_records_kernel(const int4* __restrict__ pp, int n, const int* __restrict__ rec_off,
                                    const ShellDev* __restrict__ sh, const double* __restrict__ exps,
                                    const double* __restrict__ coefs, double* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int4 q = pp[idx];
  const ShellDev A = sh[q.x], B = sh[q.y];
  const double a = exps[A.prim_off + q.z], b = exps[B.prim_off + q.w];
  const double ca = coefs[A.prim_off + q.z], cb = coefs[B.prim_off + q.w];
  const int la = A.l + 1, lb = B.l + 1;
  const double p = a + b, mu = a * b / p;
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double Ac[3] = {A.x, A.y, A.z}, Bc[3] = {B.x, B.y, B.z};
  const int NE = ne_axis(la, lb);
  double* o = out + rec_off[idx];
  o[0] = p;
  for (int d = 0; d < 3; ++d) o[1 + d] = (a * Ac[d] + b * Bc[d]) / p;
  o[4] = ca * cb * exp(-mu * (AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2]));
  o[5] = a;
  for (int d = 0; d < 3; ++d) {
    const double Q = AB[d];
    const double xpa = -b * Q / p, xpb = a * Q / p, i2p = 0.5 / p;
    double Ea[4][8];
    for (int ia = 0; ia <= la; ++ia)
      for (int t = 0; t < 8; ++t) Ea[ia][t] = 0.0;
    Ea[0][0] = 1.0;
    for (int ia = 1; ia <= la; ++ia)
      for (int t = 0; t <= ia; ++t) {
        double v = 0.0;
        if (t <= ia - 1) v += xpa * Ea[ia - 1][t];
        if (t > 0) v += Ea[ia - 1][t - 1] * i2p;
        if (t + 1 <= ia - 1) v += (t + 1) * Ea[ia - 1][t + 1];
        Ea[ia][t] = v;
      }
    for (int ia = 0; ia <= la; ++ia) {
      double cur[8], nxt[8];
      for (int t = 0; t < 8; ++t) cur[t] = Ea[ia][t];
      for (int ib = 0; ib <= lb; ++ib) {
        const int ord = ia + ib;
        const int off = 6 + d * NE + eoff(lb, ia, ib);
        for (int t = 0; t <= ord; ++t) o[off + t] = cur[t];
        if (ib == lb) break;
        for (int t = 0; t <= ord + 1; ++t) {
          double v = 0.0;
          if (t <= ord) v += xpb * cur[t];
          if (t > 0) v += cur[t - 1] * i2p;
          if (t + 1 <= ord) v += (t + 1) * cur[t + 1];
          nxt[t] = v;
        }
        for (int t = 0; t <= ord + 1; ++t) cur[t] = nxt[t];
      }
    }
  }
}

hipError_t launch_grad_records(const int4* pp, int n, const int* rec_off, const ShellDev* sh, const double* exps,
                               const double* coefs, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(grad_records_kernel, dim3((n + 127) / 128), dim3(128), 0, s, pp, n, rec_off, sh, exps, coefs,
                     out);
  return hipGetLastError();
}

// Per-atom sums of the staged 3-vectors in two fixed-order stages. Stage 1:
// one wave per chunk of at most GRAD_CHUNK slots of one atom; lanes 0..62
// read elements lane + 63 i (component lane % 3) into four partial sums,
// then component x sums lanes x, x + 3, ..., x + 60 in order.
__global__ void grad_chunk_kernel(int nchunk, const int4* __restrict__ chunks, const long long* __restrict__ base,
                                  const double* __restrict__ stage, double* __restrict__ part) {
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (c >= nchunk) return;
  const int4 ck = chunks[c];
  const double* src = stage + base[ck.x] + 3LL * ck.y;
  const int ne = 3 * ck.z;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (lane < 63) {
    int i = lane;
    for (; i + 3 * 63 < ne; i += 4 * 63) {
      a0 += src[i];
      a1 += src[i + 63];
      a2 += src[i + 2 * 63];
      a3 += src[i + 3 * 63];
    }
    for (; i < ne; i += 63) a0 += src[i];
  }
  const double v = (a0 + a1) + (a2 + a3);
  const int x = lane < 3 ? lane : 0;
  double r = 0.0;
  for (int k = 0; k < 21; ++k) r += __shfl(v, 3 * k + x, 64);
  if (lane < 3) part[3LL * c + lane] = r;
}

// Stage 2: one wave per atom over its chunks' partial sums.
__global__ void grad_reduce_kernel(int natom, const int* __restrict__ cfirst, const double* __restrict__ part,
                                   double* __restrict__ grad) {
  const int atom = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (atom >= natom) return;
  const int c0 = cfirst[atom], c1 = cfirst[atom + 1];
  for (int x = 0; x < 3; ++x) {
    double s = 0.0;
    for (int c = c0 + lane; c < c1; c += 64) s += part[3LL * c + x];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) grad[3 * atom + x] = s;
  }
}

hipError_t launch_grad_reduce(int natom, int nchunk, const int4* chunks, const int* cfirst, const long long* base,
                              const double* stage, double* part, double* grad, hipStream_t s) {
  if (natom <= 0) return hipSuccess;
  if (nchunk > 0)
    hipLaunchKernelGGL(grad_chunk_kernel, dim3((nchunk + 3) / 4), dim3(256), 0, s, nchunk, chunks, base, stage,
                       part);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((natom + 3) / 4), dim3(256), 0, s, natom, cfirst, part, grad);
  return hipGetLastError();
}

// all ordered s/p class combinations (bra class, ket class), either order
#define DFT_FOR_GRAD_COMBOS(X) \
  X(0, 0, 0, 0)                \
  X(0, 0, 1, 0)                \
  X(0, 0, 1, 1)                \
  X(1, 0, 0, 0)                \
  X(1, 0, 1, 0)                \
  X(1, 0, 1, 1)                \
  X(1, 1, 0, 0)                \
  X(1, 1, 1, 0)                \
  X(1, 1, 1, 1)

hipError_t launch_eri_grad(int la, int lb, int lc, int ld, const int2* q, const long long* slots, long long nq,
                           const int4* gpairs, const double* grec, const int* ao, const double* DJ, const double* DK0,
                           const double* DK1, int nbf, double* stage, const double* bt, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  long long g = (nq + 255) / 256;
  if (g > 256 * 16) g = 256 * 16;
  dim3 grid((unsigned)g), block(256);
#define X(A, B, C, D)                                                                                        \
  if (la == A && lb == B && lc == C && ld == D) {                                                            \
    hipLaunchKernelGGL((eri_grad_kernel<A, B, C, D>), grid, block, 0, s, q, slots, nq, gpairs, grec, ao, DJ, \
                       DK0, DK1, nbf, stage, bt);                                                            \
    return hipGetLastError();                                                                                \
  }
  DFT_FOR_GRAD_COMBOS(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace dft
