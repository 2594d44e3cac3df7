// McMurchie-Davidson building blocks for gfx950: compile-time index tables,
// the tabulated Boys function and the Hermite R_tuv recursion.
//
// Everything here is templated on the angular momenta so that every index is
// a compile-time constant after unrolling and all intermediates stay in
// VGPRs. The reference evaluates the same quantities per primitive Cartesian
// quartet with run-time loops over a dense 16 KB local box
// (shelled.py:174-348); here they are evaluated once per primitive shell
// quartet and shared by all Cartesian components.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace dft {

// ------------------------------------------------------------ static_for
template <int Begin, int End, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (Begin < End) {
    f(std::integral_constant<int, Begin>{});
    static_for<Begin + 1, End>(static_cast<F&&>(f));
  }
}

// ----------------------------------------------------- reciprocal sqrt
// FP64: hardware v_rsq_f64 estimate (measured on gfx950: 5.2e-8 relative
// over x in [e^-6, e^12]) refined by one Newton step r <- r (3 - x r^2) / 2,
// which squares the error (measured 4.3e-15 relative, ~20 ulp: far below
// what the integrals need), instead of the IEEE sqrt + divide sequences.
// FP32: the hardware estimate (~1 ulp).
__device__ __forceinline__ double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  r = r * fma(-h * r, r, 1.5);
  return r;
}
__device__ __forceinline__ float rsqrt_nr(float x) { return __builtin_amdgcn_rsqf(x); }

// -------------------------------------------------------- index algebra
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
// number of Hermite functions with t+u+v <= L
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// tetrahedral numbering of (t,u,v): by s=t+u+v, then r=u+v, then v
__host__ __device__ constexpr int hidx(int t, int u, int v) {
  int s = t + u + v, r = u + v;
  return s * (s + 1) * (s + 2) / 6 + r * (r + 1) / 2 + v;
}
__host__ __device__ constexpr int h_s(int h) { int s = 0; while (nherm(s) <= h) ++s; return s; }
__host__ __device__ constexpr int h_r(int h) {
  int s = h_s(h); int rem = h - (s == 0 ? 0 : nherm(s - 1)); int r = 0;
  while ((r + 1) * (r + 2) / 2 <= rem) ++r; return r;
}
__host__ __device__ constexpr int h_v(int h) {
  int s = h_s(h); int rem = h - (s == 0 ? 0 : nherm(s - 1)); int r = h_r(h);
  return rem - r * (r + 1) / 2;
}
__host__ __device__ constexpr int h_u(int h) { return h_r(h) - h_v(h); }
__host__ __device__ constexpr int h_t(int h) { return h_s(h) - h_r(h); }

// Cartesian component k of a shell of momentum l, in the order
// z-outer / y-middle / x-inner of shelled.py:15-23 (p: x,y,z; d: xx,xy,yy,xz,yz,zz)
__host__ __device__ constexpr int comp_z(int l, int k) {
  int z = 0; while (k >= l - z + 1) { k -= l - z + 1; ++z; } return z;
}
__host__ __device__ constexpr int comp_y(int l, int k) {
  int z = 0; while (k >= l - z + 1) { k -= l - z + 1; ++z; } return k;
}
__host__ __device__ constexpr int comp_x(int l, int k) { return l - comp_y(l, k) - comp_z(l, k); }
__host__ __device__ constexpr int comp_of(int l, int k, int axis) {
  return axis == 0 ? comp_x(l, k) : (axis == 1 ? comp_y(l, k) : comp_z(l, k));
}

// 1-D Hermite table of a primitive pair of class (la, lb): for every
// ia <= la, ib <= lb the coefficients E_t^{ia,ib}, t = 0..ia+ib, packed
// ia-major. NE = (la+1)(lb+1)(la+lb+2)/2 per axis.
__host__ __device__ constexpr int ne_axis(int la, int lb) { return (la + 1) * (lb + 1) * (la + lb + 2) / 2; }
__host__ __device__ constexpr int eoff(int lb, int ia, int ib) {
  int o = 0;
  for (int a = 0; a < ia; ++a)
    for (int b = 0; b <= lb; ++b) o += a + b + 1;
  for (int b = 0; b < ib; ++b) o += ia + b + 1;
  return o;
}
// primitive-pair record: p, Px, Py, Pz, Kab, s (Schwarz factor
// sqrt(max_x (k k|k k)_xx) of the primitive pair, 0 where not computed), then
// Ex[NE], Ey[NE], Ez[NE]
constexpr int REC_HEAD = 6;
__host__ __device__ constexpr int rec_size(int la, int lb) { return REC_HEAD + 3 * ne_axis(la, lb); }

// ------------------------------------------------------------- Boys F_m
// Table of F_m(T_k), T_k = k * BOYS_DT, k < BOYS_NG, m < BOYS_MT, generated on
// the host in extended precision. A kernel of total momentum L stages, per
// grid point k, the row F_{L+j}(T_k)/j! (j < BOYS_TAYLOR) followed by
// e^{-T_k} into LDS and evaluates
//   F_L(T) = sum_j F_{L+j}(T_k)/j! (-d)^j,  d = T - T_k,  |d| <= DT/2
// (dF_m/dT = -F_{m+1}), e^{-T} = e^{-T_k} sum_j (-d)^j/j! (eight terms,
// truncation below 1e-15 relative), then recurses downward
//   F_m = (2T F_{m+1} + e^{-T}) / (2m+1).
// For T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2 (erfc(sqrt T) < 1e-18) and the upward
// recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T), stable for m < T, with
// the e^{-T} term dropped where it is negligible (boys_tnoexp).
constexpr int BOYS_NG = 401;
constexpr double BOYS_DT = 0.1;
constexpr double BOYS_INV_DT = 10.0;
constexpr double BOYS_TMAX = 40.0;
constexpr int BOYS_TAYLOR = 8;
constexpr int BOYS_MT = 32;  // m range held in the global table
constexpr int BOYS_ROW = 10;  // LDS row: BOYS_TAYLOR coefficients, e^{-T_k}, pad (16-byte rows)
constexpr int BOYS_LDS = BOYS_NG * BOYS_ROW;  // doubles of the staged table

__host__ __device__ constexpr double inv_odd(int m) { return 1.0 / (2.0 * m + 1.0); }

// Above BOYS_TMAX, F_m = Gamma(m+1/2) / (2 T^{m+1/2}) - O(e^{-T} / (2T)). The
// e^{-T} term is dropped where it is below 1e-16 of F_L: from T = BOYS_TMAX
// for L <= 1 and from BOYS_TMAX + 4 (L - 1) above (e.g. 68 for L = 8, where it
// is needed up to T ~ 62); in between the upward recursion keeps it.
__host__ __device__ constexpr double boys_tnoexp(int L) { return L <= 1 ? BOYS_TMAX : BOYS_TMAX + 4.0 * (L - 1); }

// FP64 Boys function F_0..F_L from the LDS slice tab[k * BOYS_TAYLOR + j].
// Nearest grid point of 0 <= T < BOYS_TMAX (truncation of T/DT + 1/2) and its
// LDS row; the row offset is a 24-bit multiply (full rate, k < 2^24) rather
// than a 32-bit one.
__device__ __forceinline__ int boys_point(double T) { return (int)fma(T, BOYS_INV_DT, 0.5); }
__device__ __forceinline__ const double* boys_row(const double* tab, int k) {
  return tab + __umul24((unsigned)k, (unsigned)BOYS_ROW);
}

template <int L>
__device__ __forceinline__ void boys_fp64(double T, const double* __restrict__ tab, double (&F)[L + 1]) {
  if (T < BOYS_TMAX) {
    const int k = boys_point(T);
    double d = -(T - k * BOYS_DT);
    const double* c = boys_row(tab, k);
    double f = c[BOYS_TAYLOR - 1];
#pragma unroll
    for (int j = BOYS_TAYLOR - 2; j >= 0; --j) f = fma(f, d, c[j]);
    F[L] = f;
    if constexpr (L > 0) {
      constexpr double inv_fact[8] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
      double pe = inv_fact[7];
#pragma unroll
      for (int j = 6; j >= 0; --j) pe = fma(pe, d, inv_fact[j]);
      double e = c[BOYS_TAYLOR] * pe;
      double t2 = T + T;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], e) * inv_odd(m);
    }
  } else {
    const double rs = rsqrt_nr(T);
    F[0] = 0.88622692545275801365 * rs;  // sqrt(pi)/2 / sqrt(T)
    if constexpr (L > 0) {
      const double h = 0.5 * rs * rs;  // 1 / (2T)
      if (L <= 1 || T >= boys_tnoexp(L)) {
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = (2 * m + 1) * F[m] * h;
      } else {
        const double e = exp(-T);
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
      }
    }
  }
}

// Above boys_tasym32(L) the asymptotic form Gamma(m+1/2) / (2 T^{m+1/2})
// is exact to FP32 resolution for every m <= L (the neglected erfc / e^{-T}
// terms are below 6e-8 relative, checked against quadrature for L <= 8).
__host__ __device__ constexpr float boys_tasym32(int L) { return 16.0f + 2.5f * L; }

// FP32 Boys: below boys_tasym32(L) the table seed F_L(T_k) and its Taylor
// step stay FP64 and the downward recursion runs in FP32 with the native
// exp; above it the asymptotic form in FP32 (no table access).
template <int L>
__device__ __forceinline__ void boys_fp32(float Tf, const double* __restrict__ tab, float (&F)[L + 1]) {
  if (Tf < boys_tasym32(L)) {
    double T = (double)Tf;
    const int k = (int)fmaf(Tf, (float)BOYS_INV_DT, 0.5f);  // |d| <= DT/2 + 4e-6
    double d = -(T - k * BOYS_DT);
    const double* c = boys_row(tab, k);
    // five Taylor terms: truncation below 3e-10 relative, far under FP32 rounding
    double f = c[4];
#pragma unroll
    for (int j = 3; j >= 0; --j) f = fma(f, d, c[j]);
    F[L] = (float)f;
    if constexpr (L > 0) {
      float e = __expf(-Tf);
      float t2 = Tf + Tf;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = fmaf(t2, F[m + 1], e) * (float)inv_odd(m);
    }
  } else {
    const float rs = rsqrt_nr(Tf);
    F[0] = 0.886226925f * rs;
    if constexpr (L > 0) {
      const float h = 0.5f * rs * rs;
#pragma unroll
      for (int m = 0; m < L; ++m) F[m + 1] = (2 * m + 1) * F[m] * h;
    }
  }
}

template <int L, typename real>
__device__ __forceinline__ void boys_eval(real T, const double* __restrict__ tab, real (&F)[L + 1]) {
  if constexpr (std::is_same<real, double>::value) boys_fp64<L>(T, tab, F);
  else boys_fp32<L>(T, tab, F);
}

// Boys F_0..F_L from the global table (F_m(T_k) at BOYS_MT values per grid
// point): eight-term Taylor step at m = L, e^{-T}, downward recursion; the
// asymptotic form above BOYS_TMAX (as boys_fp64, which reads the LDS copy).
template <int L>
__device__ __forceinline__ void boys_global(double T, const double* __restrict__ g, double (&F)[L + 1]) {
  if (T < BOYS_TMAX) {
    const int k = boys_point(T);
    const double d = -(T - k * BOYS_DT);
    const double* c = g + k * BOYS_MT + L;
    constexpr double inv_fact[8] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
    double f = c[BOYS_TAYLOR - 1] * inv_fact[BOYS_TAYLOR - 1];
#pragma unroll
    for (int j = BOYS_TAYLOR - 2; j >= 0; --j) f = fma(f, d, c[j] * inv_fact[j]);
    F[L] = f;
    if constexpr (L > 0) {
      const double e = exp(-T);
      const double t2 = T + T;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], e) * inv_odd(m);
    }
  } else {
    const double rs = rsqrt_nr(T);
    F[0] = 0.88622692545275801365 * rs;
    if constexpr (L > 0) {
      const double h = 0.5 * rs * rs;
      if (L <= 1 || T >= boys_tnoexp(L)) {
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = (2 * m + 1) * F[m] * h;
      } else {
        const double e = exp(-T);
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
      }
    }
  }
}

// Stage the rows F_{L+j}(T_k)/j!, e^{-T_k} for all k into LDS (cooperative,
// whole block).
template <int L>
__device__ __forceinline__ void boys_stage(const double* __restrict__ gtab, double* __restrict__ lds) {
  constexpr double inv_fact[BOYS_TAYLOR] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
  for (int i = threadIdx.x; i < BOYS_LDS; i += blockDim.x) {
    int k = i / BOYS_ROW, j = i - k * BOYS_ROW;
    double v = 0.0;
    if (j < BOYS_TAYLOR) v = gtab[k * BOYS_MT + L + j] * inv_fact[j];
    else if (j == BOYS_TAYLOR) v = exp(-k * BOYS_DT);
    lds[i] = v;
  }
  __syncthreads();
}

// --------------------------------------------------- Hermite R^0_{tuv}
// R^n_000 = (-2a)^n F_n; R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{tuv}
// (and the same in u/Y, v/Z). Evaluated level by level from n = L down to 0
// in one register array, in place: within a level the entries are updated
// from the highest t+u+v down, so every entry still read (t+u+v lower by one
// or two) holds the previous level's value. Returns R^0 for t+u+v <= L in
// hidx order.
template <int L, typename real>
__device__ __forceinline__ void hermite_R(const real (&F)[L + 1], real m2a, real X, real Y, real Z,
                                          real (&R)[nherm(L)]) {
  real s[L + 1];
  real pw = real(1);
#pragma unroll
  for (int n = 0; n <= L; ++n) { s[n] = F[n] * pw; pw *= m2a; }
  R[0] = s[L];
  static_for<0, L>([&](auto it) {
    constexpr int n = L - 1 - decltype(it)::value;
    constexpr int S = L - n;  // this level holds t+u+v <= S
    // entries with t+u+v = k, from k = S down to 1
    static_for<0, S>([&](auto kI) {
      constexpr int k = S - decltype(kI)::value;
      static_for<nherm(k - 1), nherm(k)>([&](auto hI) {
        constexpr int h = decltype(hI)::value;
        constexpr int t = h_t(h), u = h_u(h), v = h_v(h);
        if constexpr (t > 0) {
          real r = X * R[hidx(t - 1, u, v)];
          if constexpr (t > 1) r = fma(real(t - 1), R[hidx(t - 2, u, v)], r);
          R[h] = r;
        } else if constexpr (u > 0) {
          real r = Y * R[hidx(t, u - 1, v)];
          if constexpr (u > 1) r = fma(real(u - 1), R[hidx(t, u - 2, v)], r);
          R[h] = r;
        } else {
          real r = Z * R[hidx(t, u, v - 1)];
          if constexpr (v > 1) r = fma(real(v - 1), R[hidx(t, u, v - 2)], r);
          R[h] = r;
        }
      });
    });
    R[0] = s[n];
  });
}

}  // namespace dft
