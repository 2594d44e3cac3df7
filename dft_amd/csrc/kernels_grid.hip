// Basis-function values on a grid of points: the AO evaluation of the
// reference's density post-processing (shelled_eval.py:98-113, flat variant
// main.py:905-912), one lane per grid point.
#include "kernels.h"
#include "md_device.h"

namespace dft {

// out[(ao + k) * ld + g] = phi_{ao+k}(r_g) for every shell and Cartesian
// component k: the contracted radial part sum_i c_i exp(-a_i r^2) (c_i with
// the radial normalisation, as in the integral records) times
// x^i y^j z^k / sqrt((2i-1)!!(2j-1)!!(2k-1)!!). The shell loop is
// wave-uniform, so shell data and exponents come through scalar loads; the
// point's coordinates are FP32 (the reference's grid dtype), the values FP64.
__global__ __launch_bounds__(256) void ao_grid_kernel(int nsh, const ShellDev* __restrict__ sh,
                                                      const double* __restrict__ exps,
                                                      const double* __restrict__ coefs,
                                                      const float* __restrict__ grid, long long npts,
                                                      double* __restrict__ out, long long ld) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts) return;
  const double gx = grid[3 * g], gy = grid[3 * g + 1], gz = grid[3 * g + 2];
  // 1/sqrt((2l-1)!!) for l <= 4
  const double isd[5] = {1.0, 1.0, 0.57735026918962576451, 0.25819888974716112568, 0.097590007294853320675};
  for (int s = 0; s < nsh; ++s) {
    const ShellDev S = sh[s];
    const double dx = gx - S.x, dy = gy - S.y, dz = gz - S.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double rad = 0.0;
    for (int i = 0; i < S.nprim; ++i) {
      const double a = exps[S.prim_off + i];
      // exp(-a r^2) < 1e-16 (a r^2 > 36.84) is left out: most shells are far
      // from most points of a box grid, and the exp is the kernel's main cost
      if (a * r2 < 36.84) rad += coefs[S.prim_off + i] * exp(-a * r2);
    }
    const int nc = (S.l + 1) * (S.l + 2) / 2;
    for (int k = 0; k < nc; ++k) {
      const int ix = comp_of(S.l, k, 0), iy = comp_of(S.l, k, 1), iz = comp_of(S.l, k, 2);
      double m = rad * (isd[ix] * isd[iy] * isd[iz]);
      for (int t = 0; t < ix; ++t) m *= dx;
      for (int t = 0; t < iy; ++t) m *= dy;
      for (int t = 0; t < iz; ++t) m *= dz;
      out[(long long)(S.ao + k) * ld + g] = m;
    }
  }
}

hipError_t launch_ao_grid(int nsh, const ShellDev* sh, const double* exps, const double* coefs, const float* grid,
                          long long npts, double* out, long long ld, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  const long long blocks = (npts + 255) / 256;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ao_grid_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nsh, sh, exps, coefs, grid, npts, out,
                     ld);
  return hipGetLastError();
}

}  // namespace dft
