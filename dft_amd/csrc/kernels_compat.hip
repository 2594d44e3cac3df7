// Primitive-level kernels with the data layout of the reference's two CUDA
// kernels, so callers that launch nuclear_kernel / eri_kernel themselves
// (shelled.py:1015-1036, 642-662; main.py:691-708, 719-743) can keep their
// buffers. One lane per primitive Cartesian pair / quartet, as the reference,
// but with the tabulated Boys function and a sparse R_tuv (t+u+v <= L)
// instead of the dense 16 KB box.
#include "kernels.h"
#include "md_device.h"

namespace dft {

constexpr int CMAX_L = 12;   // the reference's capacity: max_boys 13 (shelled.py:168-171)
constexpr int CMAX_E = 7;    // max_e

__device__ void boys_any(int L, double T, const double* __restrict__ g, double* F) {
  if (T < BOYS_TMAX) {
    const int k = __double2int_rn(T * BOYS_INV_DT);
    const double d = -(T - k * BOYS_DT);
    const double inv_fact[BOYS_TAYLOR] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
    double f = 0.0;
    for (int j = BOYS_TAYLOR - 1; j >= 0; --j) f = fma(f, d, g[k * BOYS_MT + L + j] * inv_fact[j]);
    F[L] = f;
    const double e = exp(-T);
    for (int m = L - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2.0 * m + 1.0);
  } else {
    F[0] = 0.88622692545275801365 / sqrt(T);
    const double e = exp(-T);
    for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

// R^0_tuv, t+u+v <= L, into R0 (hidx order), two level buffers in scratch
__device__ void hermite_R_any(int L, const double* F, double m2a, double X, double Y, double Z, double* R0) {
  double lev[2][nherm(CMAX_L)];
  double s[CMAX_L + 1], pw = 1.0;
  for (int n = 0; n <= L; ++n) { s[n] = F[n] * pw; pw *= m2a; }
  int cur = 0;
  lev[0][0] = s[L];
  for (int n = L - 1; n >= 0; --n) {
    const int S = L - n, nx = cur ^ 1;
    lev[nx][0] = s[n];
    for (int h = 1; h < nherm(S); ++h) {
      const int t = h_t(h), u = h_u(h), v = h_v(h);
      double r;
      if (t > 0) {
        r = X * lev[cur][hidx(t - 1, u, v)];
        if (t > 1) r += (t - 1) * lev[cur][hidx(t - 2, u, v)];
      } else if (u > 0) {
        r = Y * lev[cur][hidx(t, u - 1, v)];
        if (u > 1) r += (u - 1) * lev[cur][hidx(t, u - 2, v)];
      } else {
        r = Z * lev[cur][hidx(t, u, v - 1)];
        if (v > 1) r += (v - 1) * lev[cur][hidx(t, u, v - 2)];
      }
      lev[nx][h] = r;
    }
    cur = nx;
  }
  for (int h = 0; h < nherm(L); ++h) R0[h] = lev[cur][h];
}

// E_t for one axis with E_0 = exp(-ab/p Q^2): calc_E of shelled.py:228-294
__device__ void e_axis(double a, double b, double Q, int la, int lb, double* E) {
  const double p = a + b;
  double prev[CMAX_E], next[CMAX_E];
  for (int t = 0; t < CMAX_E; ++t) { prev[t] = 0.0; E[t] = 0.0; }
  prev[0] = exp(-(a * b / p) * Q * Q);
  int order = 0;
  for (int step = 0; step < la + lb; ++step) {
    const double coef = step < la ? -(b * Q / p) : (a * Q / p);
    const int no = order + 1;
    for (int t = 0; t <= no; ++t) {
      double v = 0.0;
      if (t <= order) v += coef * prev[t];
      if (t > 0) v += prev[t - 1] / (2.0 * p);
      if (t + 1 <= order) v += (t + 1) * prev[t + 1];
      next[t] = v;
    }
    order = no;
    for (int t = 0; t <= order; ++t) prev[t] = next[t];
  }
  for (int t = 0; t <= order; ++t) E[t] = prev[t];
}

__global__ void nuclear_prim_kernel(int nfa, int nfb, const double* __restrict__ ea, const double* __restrict__ eb,
                                    const double* __restrict__ ca, int sa, const double* __restrict__ cb, int sb,
                                    const int* __restrict__ pa, const int* __restrict__ pb, int natom,
                                    const int* __restrict__ Zat, const double* __restrict__ xyz,
                                    double* __restrict__ V, double* __restrict__ Ex, double* __restrict__ Ey,
                                    double* __restrict__ Ez, double* __restrict__ pp, double* __restrict__ PP,
                                    int* __restrict__ nx, int* __restrict__ ny, int* __restrict__ nz,
                                    const double* __restrict__ g) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nfa * nfb) return;
  const int a = (int)(q / nfb), b = (int)(q % nfb);
  const double e1 = ea[a], e2 = eb[b];
  const double* A = ca + (long long)sa * a;
  const double* B = cb + (long long)sb * b;
  double E[3][CMAX_E];
  int len[3];
  for (int d = 0; d < 3; ++d) {
    e_axis(e1, e2, A[d] - B[d], pa[3 * a + d], pb[3 * b + d], E[d]);
    len[d] = pa[3 * a + d] + pb[3 * b + d] + 1;
  }
  const double p = e1 + e2;
  double P[3];
  for (int d = 0; d < 3; ++d) P[d] = (e1 * A[d] + e2 * B[d]) / p;
  const int L = len[0] + len[1] + len[2] - 3;
  double value = 0.0;
  if (L <= CMAX_L) {
    for (int c = 0; c < natom; ++c) {
      const double X = P[0] - xyz[3 * c], Y = P[1] - xyz[3 * c + 1], Zc = P[2] - xyz[3 * c + 2];
      double F[CMAX_L + 1], R0[nherm(CMAX_L)];
      boys_any(L, p * (X * X + Y * Y + Zc * Zc), g, F);
      hermite_R_any(L, F, -2.0 * p, X, Y, Zc, R0);
      double s = 0.0;
      for (int t = 0; t < len[0]; ++t)
        for (int u = 0; u < len[1]; ++u)
          for (int v = 0; v < len[2]; ++v) s += E[0][t] * E[1][u] * E[2][v] * R0[hidx(t, u, v)];
      value += Zat[c] * s;
    }
  }
  V[q] = value * (-2.0 * M_PI / p);
  for (int i = 0; i < CMAX_E; ++i) {
    Ex[q * CMAX_E + i] = E[0][i];
    Ey[q * CMAX_E + i] = E[1][i];
    Ez[q * CMAX_E + i] = E[2][i];
  }
  pp[q] = p;
  for (int d = 0; d < 3; ++d) PP[3 * q + d] = P[d];
  nx[q] = len[0];
  ny[q] = len[1];
  nz[q] = len[2];
}

__global__ void eri_prim_kernel(long long nab, const double* __restrict__ pab, const double* __restrict__ Pab,
                                const double* __restrict__ Exab, const double* __restrict__ Eyab,
                                const double* __restrict__ Ezab, const int* __restrict__ nxab,
                                const int* __restrict__ nyab, const int* __restrict__ nzab, long long ncd,
                                const double* __restrict__ pcd, const double* __restrict__ Pcd,
                                const double* __restrict__ Excd, const double* __restrict__ Eycd,
                                const double* __restrict__ Ezcd, const int* __restrict__ nxcd,
                                const int* __restrict__ nycd, const int* __restrict__ nzcd, int es,
                                double* __restrict__ out, const double* __restrict__ g) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nab * ncd) return;
  const long long i = q / ncd, j = q % ncd;
  const double pi_ = pab[i], pj = pcd[j];
  const int li[3] = {nxab[i], nyab[i], nzab[i]}, lj[3] = {nxcd[j], nycd[j], nzcd[j]};
  const double* Ei[3] = {Exab + i * es, Eyab + i * es, Ezab + i * es};
  const double* Ej[3] = {Excd + j * es, Eycd + j * es, Ezcd + j * es};
  const int L = li[0] + li[1] + li[2] + lj[0] + lj[1] + lj[2] - 6;
  double val = 0.0;
  if (L <= CMAX_L) {
    const double rho = pi_ * pj / (pi_ + pj);
    const double X = Pab[3 * i] - Pcd[3 * j], Y = Pab[3 * i + 1] - Pcd[3 * j + 1], Z = Pab[3 * i + 2] - Pcd[3 * j + 2];
    double F[CMAX_L + 1], R0[nherm(CMAX_L)];
    boys_any(L, rho * (X * X + Y * Y + Z * Z), g, F);
    hermite_R_any(L, F, -2.0 * rho, X, Y, Z, R0);
    // convolution per axis (shelled.py:213-226), then the box sum over R^0
    double Cv[3][2 * CMAX_E];
    for (int d = 0; d < 3; ++d) {
      for (int s = 0; s < li[d] + lj[d] - 1; ++s) Cv[d][s] = 0.0;
      for (int t = 0; t < li[d]; ++t)
        for (int tau = 0; tau < lj[d]; ++tau) Cv[d][t + tau] += Ei[d][t] * Ej[d][tau] * ((tau & 1) ? -1.0 : 1.0);
    }
    for (int x = 0; x < li[0] + lj[0] - 1; ++x)
      for (int y = 0; y < li[1] + lj[1] - 1; ++y)
        for (int z = 0; z < li[2] + lj[2] - 1; ++z) val += Cv[0][x] * Cv[1][y] * Cv[2][z] * R0[hidx(x, y, z)];
    val *= 34.986836655249725693 / (pi_ * pj * sqrt(pi_ + pj));
  }
  out[q] = val;
}

hipError_t launch_nuclear_prim(int nfa, int nfb, const double* ea, const double* eb, const double* ca, int sa,
                               const double* cb, int sb, const int* pa, const int* pb, int natom, const int* Z,
                               const double* xyz, double* V, double* Ex, double* Ey, double* Ez, double* p,
                               double* P, int* nx, int* ny, int* nz, const double* g, hipStream_t s) {
  const long long n = (long long)nfa * nfb;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(nuclear_prim_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, nfa, nfb, ea, eb, ca, sa,
                     cb, sb, pa, pb, natom, Z, xyz, V, Ex, Ey, Ez, p, P, nx, ny, nz, g);
  return hipGetLastError();
}

hipError_t launch_eri_prim(long long nab, const double* pab, const double* Pab, const double* Exab,
                           const double* Eyab, const double* Ezab, const int* nxab, const int* nyab,
                           const int* nzab, long long ncd, const double* pcd, const double* Pcd, const double* Excd,
                           const double* Eycd, const double* Ezcd, const int* nxcd, const int* nycd,
                           const int* nzcd, int es, double* out, const double* g, hipStream_t s) {
  const long long n = nab * ncd;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(eri_prim_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, nab, pab, Pab, Exab, Eyab,
                     Ezab, nxab, nyab, nzab, ncd, pcd, Pcd, Excd, Eycd, Ezcd, nxcd, nycd, nzcd, es, out, g);
  return hipGetLastError();
}

}  // namespace dft
