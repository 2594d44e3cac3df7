// One-electron integrals S, T, V and their nuclear derivatives, templated
// per shell-pair class (la >= lb: the plan's shells are l-sorted and its
// shell pairs have the later shell first) like the ERI kernels: every index
// is a compile-time constant, so the Hermite tables stay in registers.
#include "kernels.h"
#include "md_device.h"

namespace dft {

namespace {

constexpr double PI = 3.14159265358979323846;

// 1-D Hermite coefficients of one axis, E[i][j][t] = E_t^{i,j} for
// i <= IA, j <= JB, t <= i + j (zero above), with E_0^{00} = 1: the
// recursion of calc_E (shelled.py:238-293), a raised first (X_PA), then b
// (X_PB) from each E^{i,0}.
template <int IA, int JB>
__device__ __forceinline__ void etab(double xpa, double xpb, double i2p, double (&E)[IA + 1][JB + 1][IA + JB + 1]) {
  constexpr int NT = IA + JB + 1;
#pragma unroll
  for (int i = 0; i <= IA; ++i)
#pragma unroll
    for (int j = 0; j <= JB; ++j)
#pragma unroll
      for (int t = 0; t < NT; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
#pragma unroll
  for (int i = 0; i < IA; ++i)
#pragma unroll
    for (int t = 0; t <= i + 1; ++t) {
      double v = t <= i ? xpa * E[i][0][t] : 0.0;
      if (t > 0) v = fma(i2p, E[i][0][t - 1], v);
      if (t + 1 <= i) v = fma(double(t + 1), E[i][0][t + 1], v);
      E[i + 1][0][t] = v;
    }
#pragma unroll
  for (int i = 0; i <= IA; ++i)
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = t <= i + j ? xpb * E[i][j][t] : 0.0;
        if (t > 0) v = fma(i2p, E[i][j][t - 1], v);
        if (t + 1 <= i + j) v = fma(double(t + 1), E[i][j][t + 1], v);
        E[i][j + 1][t] = v;
      }
}

// 1 / sqrt((2i-1)!!) for i <= 4
__host__ __device__ constexpr double inv_sqrt_dfac(int i) {
  return i <= 1 ? 1.0 : (i == 2 ? 0.57735026918962576451 : (i == 3 ? 0.25819888974716112568 : 0.09759000729485332159));
}
// angular normalisation of a Cartesian component pair: the part of
// shelled.py:95-101 that depends on the component
template <int LA, int LB>
__host__ __device__ constexpr double ang_norm(int ca, int cb) {
  double n = 1.0;
  for (int d = 0; d < 3; ++d) n *= inv_sqrt_dfac(comp_of(LA, ca, d)) * inv_sqrt_dfac(comp_of(LB, cb, d));
  return n;
}

// nucleus groups per primitive pair: minimises (tasks per lane) x (nuclei per task)
__device__ __forceinline__ int nucleus_groups(int npp, int nat) {
  int best = 1;
  long long cost = (long long)((npp + 63) / 64) * nat;
  for (int g = 2; g <= 64 && g <= nat; ++g) {
    const long long c = (long long)((npp * g + 63) / 64) * ((nat + g - 1) / g);
    if (c < cost) {
      cost = c;
      best = g;
    }
  }
  return best;
}

}  // namespace

// ------------------------------------------------------- S, T, V
// One 64-lane workgroup per shell pair. Lane tasks are (primitive pair k,
// nucleus group g): the nuclei of the pair's molecule are dealt round-robin
// over G groups, G chosen per pair so that the longest lane is shortest.
// The nuclear sum is taken in Hermite space (R_s = sum_C Z_C R_tuv(p, P - C)
// over the task's nuclei) and contracted once with the pair's E tables.
// Each lane accumulates into its own LDS column; the columns are summed in
// lane order at the end (results independent of timing).
//
// S from E_0 sqrt(pi/p) per axis (the binomial closed form of
// shelled.py:104-146), T from the per-axis kinetic piece of T_raw
// (shelled.py:149-155) combined as shelled.py:989-998, V as the
// nuclear_kernel sum (shelled.py:413-436): -2 pi / p sum_C Z_C
// sum_tuv E_t E_u E_v R_tuv(p, P - C). The caller's coefs carry the radial
// normalisation and the contraction coefficients (shelled.py:976-1045).
template <int LA, int LB>
__global__ __launch_bounds__(64) void one_e_kernel(const int2* __restrict__ sp, int npairs,
                                                   const ShellDev* __restrict__ sh, const double* __restrict__ exps,
                                                   const double* __restrict__ coefs, const double* __restrict__ atoms,
                                                   double* __restrict__ S, double* __restrict__ T,
                                                   double* __restrict__ V, const double* __restrict__ gtab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB, L = LA + LB, NH = nherm(L);
  __shared__ double acc[3][NAB][64];
  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= npairs) return;
  const int2 pr = sp[blockIdx.x];
  const ShellDev A = sh[pr.x], B = sh[pr.y];
#pragma unroll
  for (int e = 0; e < NAB; ++e) acc[0][e][lane] = acc[1][e][lane] = acc[2][e][lane] = 0.0;
  const int npp = A.nprim * B.nprim, nat = A.atom_end - A.atom_begin;
  const int G = nucleus_groups(npp, nat);
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  for (int t = lane; t < npp * G; t += 64) {
    const int k = t / G, grp = t - k * G;
    const int i = k / B.nprim, j = k - i * B.nprim;
    const double a = exps[A.prim_off + i], b = exps[B.prim_off + j];
    const double p = a + b, ip = 1.0 / p, i2p = 0.5 * ip;
    const double cc = coefs[A.prim_off + i] * coefs[B.prim_off + j] * exp(-a * b * ip * rab2);
    if (grp == 0) {
      // overlap and kinetic energy (b raised to lb + 2 for T)
      const double sq = sqrt(PI * ip);
      double S1[3][LA + 1][LB + 1], T1[3][LA + 1][LB + 1];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double E[LA + 1][LB + 3][L + 3];
        etab<LA, LB + 2>(-b * AB[d] * ip, a * AB[d] * ip, i2p, E);
#pragma unroll
        for (int ia = 0; ia <= LA; ++ia)
#pragma unroll
          for (int ib = 0; ib <= LB; ++ib) {
            const double s0 = E[ia][ib][0] * sq;
            double tk = fma(-2.0 * b * (2 * ib + 1), s0, 4.0 * b * b * E[ia][ib + 2][0] * sq);
            if (ib >= 2) tk = fma(double(ib * (ib - 1)), E[ia][ib - 2][0] * sq, tk);
            S1[d][ia][ib] = s0;
            T1[d][ia][ib] = tk;
          }
      }
      static_for<0, NAB>([&](auto eI) {
        constexpr int e = decltype(eI)::value, ca = e / NB, cb = e % NB;
        constexpr int ax = comp_x(LA, ca), ay = comp_y(LA, ca), az = comp_z(LA, ca);
        constexpr int bx = comp_x(LB, cb), by = comp_y(LB, cb), bz = comp_z(LB, cb);
        constexpr double an = ang_norm<LA, LB>(ca, cb);
        const double c = cc * an;
        const double sx = S1[0][ax][bx], sy = S1[1][ay][by], sz = S1[2][az][bz];
        acc[0][e][lane] += c * sx * sy * sz;
        acc[1][e][lane] += c * -0.5 * (T1[0][ax][bx] * sy * sz + sx * T1[1][ay][by] * sz + sx * sy * T1[2][az][bz]);
      });
    }
    // nuclear attraction: Hermite sum over the task's nuclei, then E E E
    const double P[3] = {(a * A.x + b * B.x) * ip, (a * A.y + b * B.y) * ip, (a * A.z + b * B.z) * ip};
    double Rs[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) Rs[h] = 0.0;
    for (int c = A.atom_begin + grp; c < A.atom_end; c += G) {
      const double X = P[0] - atoms[4 * c], Y = P[1] - atoms[4 * c + 1], Z = P[2] - atoms[4 * c + 2];
      const double Zc = atoms[4 * c + 3];
      double F[L + 1], R[NH];
      boys_global<L>(p * (X * X + Y * Y + Z * Z), gtab, F);
      hermite_R<L, double>(F, -2.0 * p, X, Y, Z, R);
#pragma unroll
      for (int h = 0; h < NH; ++h) Rs[h] = fma(Zc, R[h], Rs[h]);
    }
    double Ev[3][LA + 1][LB + 1][L + 1];
#pragma unroll
    for (int d = 0; d < 3; ++d) etab<LA, LB>(-b * AB[d] * ip, a * AB[d] * ip, i2p, Ev[d]);
    const double pre = cc * (-2.0 * PI * ip);
    static_for<0, NAB>([&](auto eI) {
      constexpr int e = decltype(eI)::value, ca = e / NB, cb = e % NB;
      constexpr int ax = comp_x(LA, ca), ay = comp_y(LA, ca), az = comp_z(LA, ca);
      constexpr int bx = comp_x(LB, cb), by = comp_y(LB, cb), bz = comp_z(LB, cb);
      constexpr double an = ang_norm<LA, LB>(ca, cb);
      double v = 0.0;
#pragma unroll
      for (int tt = 0; tt <= ax + bx; ++tt)
#pragma unroll
        for (int uu = 0; uu <= ay + by; ++uu) {
          double w = 0.0;
#pragma unroll
          for (int vv = 0; vv <= az + bz; ++vv) w = fma(Ev[2][az][bz][vv], Rs[hidx(tt, uu, vv)], w);
          v = fma(Ev[0][ax][bx][tt] * Ev[1][ay][by][uu], w, v);
        }
      acc[2][e][lane] += pre * an * v;
    });
  }
  __syncthreads();
  for (int e = lane; e < 3 * NAB; e += 64) {
    const int kind = e / NAB, ab = e - kind * NAB, ca = ab / NB, cb = ab - ca * NB;
    // a diagonal shell pair writes each element pair once, from ca >= cb
    if (pr.x == pr.y && ca < cb) continue;
    double s = 0.0;
    for (int l = 0; l < 64; ++l) s += acc[kind][ab][l];
    const long long r = A.ao + ca, c = B.ao + cb;
    const long long nbf = A.nbf, rc = A.mbase + r * nbf + c, cr = A.mbase + c * nbf + r;
    double* out = kind == 0 ? S : (kind == 1 ? T : V);
    out[rc] = s;
    out[cr] = s;
  }
}

// ------------------------------------------ 1-e nuclear derivatives
// Per shell pair (one molecule, natom nuclei at atoms[0..natom)), with the
// density P and the energy-weighted density W (symmetric, nbf x nbf): the
// derivatives of sum_ab P_ab (T_ab + V_ab) - sum_ab W_ab S_ab with respect
// to the centre of A, of B and of each nucleus C.
// Phase 1, lane per primitive pair (at most G1_PPMAX per pass): the S and T
// derivatives with respect to A (B's are their negative) and the Hermite
// vectors D^{X,d}_tuv = sum_ab P_ab c_ab d/dX_d (E_t E_u E_v) for X = A, B,
// from dE^{ij}_t/dA = 2a E^{i+1,j}_t - i E^{i-1,j}_t (b and j for B), kept
// in LDS. Phase 2, lane per nucleus (64 per chunk): R_tuv(p, P - C) up to
// L + 1 and six dot products per primitive pair. The nucleus part is minus
// the sum of the A and B parts (V_ab for a fixed C is invariant under a
// common translation of A, B and C). Each (shell pair, nucleus) writes
// three 3-vectors (atom of A, atom of B, the nucleus) to its slots; the S/T
// part goes with nucleus 0.
constexpr int G1_PPMAX = 36;  // primitive pairs per LDS pass
template <int LA, int LB>
__global__ __launch_bounds__(64) void grad1e_kernel(const int2* __restrict__ sp, const int* __restrict__ pidx,
                                                    int npairs, const ShellDev* __restrict__ sh,
                                                    const double* __restrict__ exps, const double* __restrict__ coefs,
                                                    int natom, const double* __restrict__ atoms,
                                                    const double* __restrict__ Pm, const double* __restrict__ Wm,
                                                    int nbf, const long long* __restrict__ slots,
                                                    double* __restrict__ stage, const double* __restrict__ gtab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), L1 = LA + LB + 1, NH = nherm(L1);
  __shared__ double D[G1_PPMAX][6][NH];
  __shared__ double head[G1_PPMAX][4];  // p, Px, Py, Pz
  __shared__ double st[64][3];          // S/T derivatives (A centre) per lane
  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= npairs) return;
  const int kpair = pidx[blockIdx.x];
  const int2 pr = sp[kpair];
  const ShellDev A = sh[pr.x], B = sh[pr.y];
  // off-diagonal shell pairs stand for both (ab) and (ba)
  const double sym = pr.x == pr.y ? 1.0 : 2.0;
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const int npp = A.nprim * B.nprim;
  double acc_st[3] = {0.0, 0.0, 0.0};
  for (int c0 = 0; c0 < natom; c0 += 64) {
    const int cn = c0 + lane;
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < npp; k0 += G1_PPMAX) {
      const int nk = min(G1_PPMAX, npp - k0);
      __syncthreads();
      if (lane < nk) {
        const int k = k0 + lane, i = k / B.nprim, j = k - i * B.nprim;
        const double a = exps[A.prim_off + i], b = exps[B.prim_off + j];
        const double p = a + b, ip = 1.0 / p, i2p = 0.5 * ip;
        const double cc = coefs[A.prim_off + i] * coefs[B.prim_off + j] * exp(-a * b * ip * rab2);
        double E[3][LA + 2][LB + 4][LA + LB + 5];
#pragma unroll
        for (int d = 0; d < 3; ++d) etab<LA + 1, LB + 3>(-b * AB[d] * ip, a * AB[d] * ip, i2p, E[d]);
        const double sq = sqrt(PI * ip);
        double Dl[6][NH];
#pragma unroll
        for (int x = 0; x < 6; ++x)
#pragma unroll
          for (int h = 0; h < NH; ++h) Dl[x][h] = 0.0;
        static_for<0, NA * NB>([&](auto eI) {
          constexpr int e = decltype(eI)::value, ca = e / NB, cb = e % NB;
          constexpr int pa[3] = {comp_x(LA, ca), comp_y(LA, ca), comp_z(LA, ca)};
          constexpr int pb[3] = {comp_x(LB, cb), comp_y(LB, cb), comp_z(LB, cb)};
          constexpr double an = ang_norm<LA, LB>(ca, cb);
          const long long r = A.ao + ca, c = B.ao + cb;
          const double Pab = sym * Pm[r * nbf + c] * cc * an;
          if (c0 == 0) {
            const double Wab = sym * Wm[r * nbf + c] * cc * an;
            // 1-D overlap S1 and kinetic piece T1 (shelled.py:149-155)
            auto S1 = [&](int d, int ia, int ib) { return E[d][ia][ib][0] * sq; };
            auto T1 = [&](int d, int ia, int ib) {
              double t = fma(-2.0 * b * (2 * ib + 1), S1(d, ia, ib), 4.0 * b * b * S1(d, ia, ib + 2));
              if (ib >= 2) t = fma(double(ib * (ib - 1)), S1(d, ia, ib - 2), t);
              return t;
            };
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
              double dS = 2.0 * a * S1(d, pa[d] + 1, pb[d]);
              double dT = 2.0 * a * T1(d, pa[d] + 1, pb[d]);
              if (pa[d] > 0) {
                dS = fma(-double(pa[d]), S1(d, pa[d] - 1, pb[d]), dS);
                dT = fma(-double(pa[d]), T1(d, pa[d] - 1, pb[d]), dT);
              }
              const double s1 = S1(o1, pa[o1], pb[o1]), s2 = S1(o2, pa[o2], pb[o2]);
              const double gS = dS * s1 * s2;
              const double gT =
                  -0.5 * (dT * s1 * s2 + dS * (T1(o1, pa[o1], pb[o1]) * s2 + s1 * T1(o2, pa[o2], pb[o2])));
              acc_st[d] += Pab * gT - Wab * gS;
            }
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
#pragma unroll
            for (int tt = 0; tt <= pa[d] + pb[d] + 1; ++tt) {
              double dA = 2.0 * a * E[d][pa[d] + 1][pb[d]][tt];
              if (pa[d] > 0) dA = fma(-double(pa[d]), E[d][pa[d] - 1][pb[d]][tt], dA);
              double dB = 2.0 * b * E[d][pa[d]][pb[d] + 1][tt];
              if (pb[d] > 0) dB = fma(-double(pb[d]), E[d][pa[d]][pb[d] - 1][tt], dB);
#pragma unroll
              for (int uu = 0; uu <= pa[o1] + pb[o1]; ++uu)
#pragma unroll
                for (int vv = 0; vv <= pa[o2] + pb[o2]; ++vv) {
                  int hh[3] = {0, 0, 0};
                  hh[d] = tt;
                  hh[o1] = uu;
                  hh[o2] = vv;
                  const int h = hidx(hh[0], hh[1], hh[2]);
                  const double ev = E[o1][pa[o1]][pb[o1]][uu] * E[o2][pa[o2]][pb[o2]][vv];
                  Dl[d][h] = fma(Pab * dA, ev, Dl[d][h]);
                  Dl[3 + d][h] = fma(Pab * dB, ev, Dl[3 + d][h]);
                }
            }
          }
        });
        const double pre = -2.0 * PI * ip;
#pragma unroll
        for (int x = 0; x < 6; ++x)
#pragma unroll
          for (int h = 0; h < NH; ++h) D[lane][x][h] = pre * Dl[x][h];
        head[lane][0] = p;
        head[lane][1] = (a * A.x + b * B.x) * ip;
        head[lane][2] = (a * A.y + b * B.y) * ip;
        head[lane][3] = (a * A.z + b * B.z) * ip;
      }
      __syncthreads();
      if (cn < natom) {
        const double Cx = atoms[4 * cn], Cy = atoms[4 * cn + 1], Cz = atoms[4 * cn + 2], Zc = atoms[4 * cn + 3];
        for (int k = 0; k < nk; ++k) {
          const double p = head[k][0];
          const double X = head[k][1] - Cx, Y = head[k][2] - Cy, Z = head[k][3] - Cz;
          double F[L1 + 1], R[NH];
          boys_global<L1>(p * (X * X + Y * Y + Z * Z), gtab, F);
          hermite_R<L1, double>(F, -2.0 * p, X, Y, Z, R);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double sa = 0.0, sb = 0.0;
#pragma unroll
            for (int h = 0; h < NH; ++h) {
              sa = fma(D[k][d][h], R[h], sa);
              sb = fma(D[k][3 + d][h], R[h], sb);
            }
            gA[d] = fma(Zc, sa, gA[d]);
            gB[d] = fma(Zc, sb, gB[d]);
          }
        }
      }
    }
    // the S/T parts: fixed-order sum over the primitive lanes, added to nucleus 0
    double tst[3] = {0.0, 0.0, 0.0};
    if (c0 == 0) {
      __syncthreads();
#pragma unroll
      for (int d = 0; d < 3; ++d) st[lane][d] = acc_st[d];
      __syncthreads();
      if (lane == 0)
        for (int l = 0; l < 64; ++l)
#pragma unroll
          for (int d = 0; d < 3; ++d) tst[d] += st[l][d];
    }
    if (cn < natom) {
      const long long* sl = slots + 3 * ((long long)kpair * natom + cn);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        stage[sl[0] + d] = gA[d] + tst[d];
        stage[sl[1] + d] = gB[d] - tst[d];
        stage[sl[2] + d] = -(gA[d] + gB[d]);
      }
    }
  }
}

// ------------------------------------- S, T, V of f-shell pairs
// The same integrals as one_e_kernel for shell pairs with an f shell, with
// the momenta at run time (no compile-time instantiation, whose register
// tables would not fit): one 64-lane workgroup per shell pair walks its
// primitive pairs in turn. Per primitive pair, three lanes build the axes'
// Hermite tables (b raised to lb + 2 for T) in LDS, every lane evaluates
// R_tuv(p, P - C) for its nuclei in its own LDS row and sums them with the
// charges, the rows are summed in lane order, and each lane adds the
// primitive pair's S, T, V to the component pairs it owns (ab = lane,
// lane + 64).
constexpr int OE_GLMAX = 3;                                 // largest momentum
constexpr int OE_GL = 2 * OE_GLMAX;                         // L of the pair
constexpr int OE_GNH = nherm(OE_GL);                        // 84
constexpr int OE_GET = (OE_GLMAX + 1) * (OE_GLMAX + 3) * (OE_GL + 3);  // one axis table
constexpr int OE_GNAB = ncart(OE_GLMAX) * ncart(OE_GLMAX);  // 100

__global__ __launch_bounds__(64) void one_e_generic_kernel(int la, int lb, const int2* __restrict__ sp, int npairs,
                                                           const ShellDev* __restrict__ sh,
                                                           const double* __restrict__ exps,
                                                           const double* __restrict__ coefs,
                                                           const double* __restrict__ atoms, double* __restrict__ S,
                                                           double* __restrict__ T, double* __restrict__ V,
                                                           const double* __restrict__ gtab) {
  __shared__ double Et[3][OE_GET];          // E[i][j][t], i <= la, j <= lb + 2, t <= i + j
  __shared__ double Rl[64][OE_GNH + 8];     // per lane: R row, then F
  __shared__ double Rs[OE_GNH];
  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= npairs) return;
  const int2 pr = sp[blockIdx.x];
  const ShellDev A = sh[pr.x], B = sh[pr.y];
  const int L = la + lb, nh = nherm(L), na = ncart(la), nb = ncart(lb), nab = na * nb;
  const int JB = lb + 2, NT = la + JB + 1;  // table extents
  auto eidx = [&](int i, int j, int t) { return (i * (JB + 1) + j) * NT + t; };
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  double accS[2] = {0.0, 0.0}, accT[2] = {0.0, 0.0}, accV[2] = {0.0, 0.0};
  double* R = Rl[lane];
  double* F = Rl[lane] + OE_GNH;
  for (int k = 0; k < A.nprim * B.nprim; ++k) {
    const int i = k / B.nprim, j = k - i * B.nprim;
    const double a = exps[A.prim_off + i], b = exps[B.prim_off + j];
    const double p = a + b, ip = 1.0 / p, i2p = 0.5 * ip;
    const double cc = coefs[A.prim_off + i] * coefs[B.prim_off + j] * exp(-a * b * ip * rab2);
    const double P[3] = {(a * A.x + b * B.x) * ip, (a * A.y + b * B.y) * ip, (a * A.z + b * B.z) * ip};
    __syncthreads();  // the previous primitive pair's readers are done
    if (lane < 3) {
      // calc_E's recursion (shelled.py:238-293) along axis `lane`
      double* E = Et[lane];
      const double xpa = -b * AB[lane] * ip, xpb = a * AB[lane] * ip;
      for (int x = 0; x < (la + 1) * (JB + 1) * NT; ++x) E[x] = 0.0;
      E[eidx(0, 0, 0)] = 1.0;
      for (int ia = 0; ia < la; ++ia)
        for (int t = 0; t <= ia + 1; ++t) {
          double v = t <= ia ? xpa * E[eidx(ia, 0, t)] : 0.0;
          if (t > 0) v = fma(i2p, E[eidx(ia, 0, t - 1)], v);
          if (t + 1 <= ia) v = fma(double(t + 1), E[eidx(ia, 0, t + 1)], v);
          E[eidx(ia + 1, 0, t)] = v;
        }
      for (int ia = 0; ia <= la; ++ia)
        for (int jb = 0; jb < JB; ++jb)
          for (int t = 0; t <= ia + jb + 1; ++t) {
            double v = t <= ia + jb ? xpb * E[eidx(ia, jb, t)] : 0.0;
            if (t > 0) v = fma(i2p, E[eidx(ia, jb, t - 1)], v);
            if (t + 1 <= ia + jb) v = fma(double(t + 1), E[eidx(ia, jb, t + 1)], v);
            E[eidx(ia, jb + 1, t)] = v;
          }
    }
    // this lane's nuclei: R summed with the charges (in place, level by level)
    for (int h = 0; h < nh; ++h) R[h] = 0.0;
    for (int c = A.atom_begin + lane; c < A.atom_end; c += 64) {
      const double X = P[0] - atoms[4 * c], Y = P[1] - atoms[4 * c + 1], Z = P[2] - atoms[4 * c + 2];
      const double Zc = atoms[4 * c + 3];
      const double Tb = p * (X * X + Y * Y + Z * Z);
      // F_0..F_L as boys_global, run-time L
      if (Tb < BOYS_TMAX) {
        const int kk = boys_point(Tb);
        const double d = -(Tb - kk * BOYS_DT);
        const double inv_fact[BOYS_TAYLOR] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
        const double* row = gtab + kk * BOYS_MT + L;
        double f = row[BOYS_TAYLOR - 1] * inv_fact[BOYS_TAYLOR - 1];
        for (int jj = BOYS_TAYLOR - 2; jj >= 0; --jj) f = fma(f, d, row[jj] * inv_fact[jj]);
        F[L] = f;
        const double e = exp(-Tb), t2 = Tb + Tb;
        for (int m = L - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], e) * inv_odd(m);
      } else {
        const double r = 1.0 / sqrt(Tb), hh = 0.5 * r * r;
        const double e = Tb >= boys_tnoexp(L) ? 0.0 : exp(-Tb);
        F[0] = 0.88622692545275801365 * r;
        for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * hh;
      }
      // R^n: level n = L holds R_000 only; each lower level updates the
      // entries with t+u+v = k from k = L - n down to 1 (in place: they read
      // entries one or two lower, still of the level above), then R_000
      double pw = 1.0;
      for (int m = 0; m < L; ++m) pw *= -2.0 * p;
      double Rt[OE_GNH];  // run-time indexed: private (scratch) memory
      Rt[0] = F[L] * pw;
      for (int n = L - 1; n >= 0; --n) {
        for (int kk = L - n; kk >= 1; --kk)
          for (int h = nherm(kk - 1); h < nherm(kk); ++h) {
            const int t = h_t(h), u = h_u(h), v = h_v(h);
            double r;
            if (t > 0) {
              r = X * Rt[hidx(t - 1, u, v)];
              if (t > 1) r = fma(double(t - 1), Rt[hidx(t - 2, u, v)], r);
            } else if (u > 0) {
              r = Y * Rt[hidx(t, u - 1, v)];
              if (u > 1) r = fma(double(u - 1), Rt[hidx(t, u - 2, v)], r);
            } else {
              r = Z * Rt[hidx(t, u, v - 1)];
              if (v > 1) r = fma(double(v - 1), Rt[hidx(t, u, v - 2)], r);
            }
            Rt[h] = r;
          }
        pw /= -2.0 * p;
        Rt[0] = F[n] * pw;
      }
      for (int h = 0; h < nh; ++h) R[h] = fma(Zc, Rt[h], R[h]);
    }
    __syncthreads();
    for (int h = lane; h < nh; h += 64) {
      double x = 0.0;
      for (int l = 0; l < 64; ++l) x += Rl[l][h];
      Rs[h] = x;
    }
    __syncthreads();
    const double sq = sqrt(PI * ip), pre = cc * (-2.0 * PI * ip);
    for (int u2 = 0; u2 < 2; ++u2) {
      const int ab = lane + 64 * u2;
      if (ab >= nab) break;
      const int ca = ab / nb, cb = ab - ca * nb;
      int pa[3], pb[3];
      double n = 1.0;
      for (int d = 0; d < 3; ++d) {
        pa[d] = comp_of(la, ca, d);
        pb[d] = comp_of(lb, cb, d);
        n *= inv_sqrt_dfac(pa[d]) * inv_sqrt_dfac(pb[d]);
      }
      double S1[3], T1[3];
      for (int d = 0; d < 3; ++d) {
        const double* E = Et[d];
        const double s0 = E[eidx(pa[d], pb[d], 0)] * sq;
        double tk = fma(-2.0 * b * (2 * pb[d] + 1), s0, 4.0 * b * b * E[eidx(pa[d], pb[d] + 2, 0)] * sq);
        if (pb[d] >= 2) tk = fma(double(pb[d] * (pb[d] - 1)), E[eidx(pa[d], pb[d] - 2, 0)] * sq, tk);
        S1[d] = s0;
        T1[d] = tk;
      }
      accS[u2] += cc * n * S1[0] * S1[1] * S1[2];
      accT[u2] += cc * n * -0.5 * (T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2]);
      double v = 0.0;
      for (int tt = 0; tt <= pa[0] + pb[0]; ++tt)
        for (int uu = 0; uu <= pa[1] + pb[1]; ++uu) {
          double w = 0.0;
          for (int vv = 0; vv <= pa[2] + pb[2]; ++vv)
            w = fma(Et[2][eidx(pa[2], pb[2], vv)], Rs[hidx(tt, uu, vv)], w);
          v = fma(Et[0][eidx(pa[0], pb[0], tt)] * Et[1][eidx(pa[1], pb[1], uu)], w, v);
        }
      accV[u2] += pre * n * v;
    }
  }
  for (int u2 = 0; u2 < 2; ++u2) {
    const int ab = lane + 64 * u2;
    if (ab >= nab) break;
    const int ca = ab / nb, cb = ab - ca * nb;
    if (pr.x == pr.y && ca < cb) continue;
    const long long r = A.ao + ca, c = B.ao + cb;
    const long long nbf = A.nbf, rc = A.mbase + r * nbf + c, cr = A.mbase + c * nbf + r;
    S[rc] = S[cr] = accS[u2];
    T[rc] = T[cr] = accT[u2];
    V[rc] = V[cr] = accV[u2];
  }
}

namespace {

template <int LA, int LB>
hipError_t launch_one_e_class(const int2* sp, int n, const ShellDev* sh, const double* exps, const double* coefs,
                              const double* atoms, double* S, double* T, double* V, const double* gtab,
                              hipStream_t s) {
  hipLaunchKernelGGL((one_e_kernel<LA, LB>), dim3(n), dim3(64), 0, s, sp, n, sh, exps, coefs, atoms, S, T, V, gtab);
  return hipGetLastError();
}

template <int LA, int LB>
hipError_t launch_grad1e_class(const int2* sp, const int* pidx, int n, const ShellDev* sh, const double* exps,
                               const double* coefs, int natom, const double* atoms, const double* P,
                               const double* W, int nbf, const long long* slots, double* stage, const double* gtab,
                               hipStream_t s) {
  hipLaunchKernelGGL((grad1e_kernel<LA, LB>), dim3(n), dim3(64), 0, s, sp, pidx, n, sh, exps, coefs, natom, atoms, P,
                     W, nbf, slots, stage, gtab);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_one_electron(int la, int lb, const int2* sp, int npairs, const ShellDev* sh, const double* exps,
                               const double* coefs, const double* atoms, double* S, double* T, double* V,
                               const double* gtab, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  if (la < lb || lb < 0 || la > OE_GLMAX) return hipErrorInvalidValue;
  if (la > 2) {
    hipLaunchKernelGGL(one_e_generic_kernel, dim3(npairs), dim3(64), 0, s, la, lb, sp, npairs, sh, exps, coefs, atoms,
                       S, T, V, gtab);
    return hipGetLastError();
  }
  hipError_t e = hipErrorInvalidValue;
  static_for<0, 3>([&](auto aI) {
    constexpr int A = decltype(aI)::value;
    static_for<0, A + 1>([&](auto bI) {
      constexpr int B = decltype(bI)::value;
      if (la == A && lb == B) e = launch_one_e_class<A, B>(sp, npairs, sh, exps, coefs, atoms, S, T, V, gtab, s);
    });
  });
  return e;
}

hipError_t launch_grad1e(int la, int lb, const int2* sp, const int* pidx, int npairs, const ShellDev* sh,
                         const double* exps, const double* coefs, int natom, const double* atoms, const double* P,
                         const double* W, int nbf, const long long* slots, double* stage, const double* gtab,
                         hipStream_t s) {
  if (npairs <= 0 || natom <= 0) return hipSuccess;
  hipError_t e = hipErrorInvalidValue;
  static_for<0, GRAD_LMAX + 1>([&](auto aI) {
    constexpr int A = decltype(aI)::value;
    static_for<0, A + 1>([&](auto bI) {
      constexpr int B = decltype(bI)::value;
      if (la == A && lb == B)
        e = launch_grad1e_class<A, B>(sp, pidx, npairs, sh, exps, coefs, natom, atoms, P, W, nbf, slots, stage, gtab,
                                      s);
    });
  });
  return e;
}

}  // namespace dft
