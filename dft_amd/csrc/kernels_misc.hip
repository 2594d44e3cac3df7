// Pair pre-pass, fixed-order reduction, symmetrisation, one-electron
// integrals and the Boys test hook.
#include "kernels.h"
#include "md_device.h"

namespace dft {

// ------------------------------------------------------- pair pre-pass
// One lane per primitive pair (i_a of shell a, i_b of shell b). Record:
//   p, Px, Py, Pz, K_ab = c_a c_b exp(-ab/p |A-B|^2) / p, the Schwarz factor
//   sq[idx] (0 without sq), then per axis the
//   table E_t^{ia,ib} (ia <= la, ib <= lb) of calc_E (shelled.py:228-294),
//   with E_0^{00} = 1 and the angular factor 1/sqrt((2ia-1)!!(2ib-1)!!).
// a is raised first (X_PA = -b Q/p), then b (X_PB = a Q/p), as in calc_E.
__global__ void pair_records_kernel(const int4* __restrict__ pp, int n, const int* __restrict__ rec_off,
                                    const ShellDev* __restrict__ sh, const double* __restrict__ exps,
                                    const double* __restrict__ coefs, const double* __restrict__ sq,
                                    double* __restrict__ r64, float* __restrict__ r32) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int4 q = pp[idx];
  const ShellDev A = sh[q.x], B = sh[q.y];
  const double a = exps[A.prim_off + q.z], b = exps[B.prim_off + q.w];
  const double ca = coefs[A.prim_off + q.z], cb = coefs[B.prim_off + q.w];
  const int la = A.l, lb = B.l;
  const double p = a + b, mu = a * b / p;
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double Ac[3] = {A.x, A.y, A.z}, Bc[3] = {B.x, B.y, B.z};
  const int NE = ne_axis(la, lb);
  // either output may be absent (FP64 and FP32 records have separate lists)
  double* o64 = r64 ? r64 + rec_off[idx] : nullptr;
  float* o32 = r32 ? r32 + rec_off[idx] : nullptr;
  double head[REC_HEAD];
  for (int d = 0; d < 3; ++d) head[1 + d] = (a * Ac[d] + b * Bc[d]) / p;
  head[0] = p;
  head[4] = ca * cb * exp(-mu * (AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2])) / p;
  head[5] = sq ? sq[idx] : 0.0;
  for (int k = 0; k < REC_HEAD; ++k) {
    if (o64) o64[k] = head[k];
    if (o32) o32[k] = (float)head[k];
  }
  // (2l-1)!! for l <= 4: 1, 1, 3, 15, 105
  const double dfac[5] = {1.0, 1.0, 3.0, 15.0, 105.0};
  for (int d = 0; d < 3; ++d) {
    const double Q = AB[d];
    const double xpa = -b * Q / p, xpb = a * Q / p, i2p = 0.5 / p;
    double Ea[5][9];  // E^{ia,0}_t
    for (int ia = 0; ia <= la; ++ia)
      for (int t = 0; t < 9; ++t) Ea[ia][t] = 0.0;
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
      double cur[9], nxt[9];
      for (int t = 0; t < 9; ++t) cur[t] = Ea[ia][t];
      for (int ib = 0; ib <= lb; ++ib) {
        const int ord = ia + ib;
        const double nrm = 1.0 / sqrt(dfac[ia] * dfac[ib]);
        const int o = REC_HEAD + d * NE + eoff(lb, ia, ib);
        for (int t = 0; t <= ord; ++t) {
          if (o64) o64[o + t] = cur[t] * nrm;
          if (o32) o32[o + t] = (float)(cur[t] * nrm);
        }
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

hipError_t launch_pair_records(const int4* pp, int n, const int* rec_off, const ShellDev* sh,
                               const double* exps, const double* coefs, const double* sq, double* r64,
                               float* r32, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(pair_records_kernel, dim3((n + 127) / 128), dim3(128), 0, s, pp, n, rec_off, sh, exps,
                     coefs, sq, r64, r32);
  return hipGetLastError();
}

__global__ void interleave_records_kernel(const int4* __restrict__ map, int nrec, const double* __restrict__ r64,
                                          const float* __restrict__ r32, double* __restrict__ t64,
                                          float* __restrict__ t32) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const int4 m = map[r];
  for (int f = 0; f < m.z; ++f) {
    if (t64) t64[m.y + (long long)f * m.w] = r64[m.x + f];
    if (t32) t32[m.y + (long long)f * m.w] = r32[m.x + f];
  }
}

hipError_t launch_interleave_records(const int4* map, int nrec, const double* r64, const float* r32, double* t64,
                                     float* t32, hipStream_t s) {
  if (nrec <= 0) return hipSuccess;
  hipLaunchKernelGGL(interleave_records_kernel, dim3((nrec + 127) / 128), dim3(128), 0, s, map, nrec, r64, r32,
                     t64, t32);
  return hipGetLastError();
}

// ------------------------------------------------- fixed-order reduction
// The half matrices are never formed: one 64-lane wave per unordered shell
// pair {X, Y} of one matrix (J or a K) reduces the contribution blocks of
// its two targets, (X, Y) and (Y, X) (one target when X = Y), and writes
// both orientations of J = Jh + Jh^T (likewise K): element (x, y) of the
// output is S_XY[x][y] + S_YX[y][x], the sum the separate symmetrisation
// pass formed from the half matrix (the same additions, so the same bits).
//
// A target's cnt contribution blocks (ne = nrow * ncol elements each) are
// contiguous in the staging buffer, so the wave streams them as one
// contiguous array: with g = 64 / ne groups, lane l < g * ne takes element
// l % ne of blocks l / ne, l / ne + g, ... (consecutive lanes read
// consecutive addresses; RED_ACC partial sums per lane). The groups are then
// combined in a fixed order (a xor butterfly when ne = 1, else group 0, 1,
// ... by lane shuffles), so J and K are bitwise reproducible run to run (the
// reference's add.at scatter is not: shelled.py:857-859). Blocks wider than
// a wave (f x f) sum element by element. The two targets' sums go through
// LDS to meet their transposes.
constexpr int RED_ACC = 16;
constexpr int RED_NE_MAX = ncart(LMAX) * ncart(LMAX);

// Sum of target t's blocks into dst[e], e < nrow * ncol (one wave).
__device__ __forceinline__ void reduce_target(const Target& t, const double* __restrict__ src, int cnt, int lane,
                                              double* __restrict__ dst) {
  const int ne = t.nrow * t.ncol;
  if (ne > 64) {
    for (int e = lane; e < ne; e += 64) {
      double acc[RED_ACC];
#pragma unroll
      for (int k = 0; k < RED_ACC; ++k) acc[k] = 0.0;
      int k = 0;
      for (; k + RED_ACC <= cnt; k += RED_ACC) {
#pragma unroll
        for (int j = 0; j < RED_ACC; ++j) acc[j] += src[(long long)(k + j) * ne + e];
      }
      for (; k < cnt; ++k) acc[0] += src[(long long)k * ne + e];
#pragma unroll
      for (int w = 1; w < RED_ACC; w *= 2)
#pragma unroll
        for (int j = 0; j + w < RED_ACC; j += 2 * w) acc[j] += acc[j + w];
      dst[e] = acc[0];
    }
    return;
  }
  const long long tot = (long long)cnt * ne;
  const int g = 64 / ne;
  const int span = g * ne;
  // RED_ACC independent partial sums keep that many loads in flight per
  // lane (the stream is latency-bound per wave: 4, 8, 16 partial sums read
  // a 168 MB staging buffer in 67, 47, 42 us)
  double acc[RED_ACC];
#pragma unroll
  for (int k = 0; k < RED_ACC; ++k) acc[k] = 0.0;
  if (lane < span) {
    long long i = lane;
    for (; i + (RED_ACC - 1) * span < tot; i += RED_ACC * span) {
#pragma unroll
      for (int k = 0; k < RED_ACC; ++k) acc[k] += src[i + k * span];
    }
    for (; i < tot; i += span) acc[0] += src[i];
  }
  // pairwise, fixed order
#pragma unroll
  for (int w = 1; w < RED_ACC; w *= 2)
#pragma unroll
    for (int k = 0; k + w < RED_ACC; k += 2 * w) acc[k] += acc[k + w];
  double s = acc[0];
  if (ne == 1) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) dst[0] = s;
  } else {
    const int el = lane < ne ? lane : 0;
    double r = 0.0;
    for (int k = 0; k < g; ++k) r += __shfl(s, k * ne + el, 64);
    if (lane < ne) dst[lane] = r;
  }
}

//
// Two register layouts of the same sums. UNIFORM makes the wave index
// explicitly wave-uniform, so the targets' base pointers sit in scalar
// registers (fewer VGPRs, more waves per SIMD): faster when the staging
// buffer streams from HBM (64 uneven (H2O)_n clusters, 7.2 GB: batch build
// 15.71 -> 15.48 ms), slower when it fits in the 256 MB Infinity Cache
// ((H2O)16, 168 MB: 38 -> 49 us), where the extra waves in flight cost more
// than they hide.
template <bool UNIFORM>
__global__ __launch_bounds__(64) void reduce_sym_kernel(int npair, const int2* __restrict__ tpairs,
                                                        const Target* __restrict__ tg,
                                                        const long long* __restrict__ tbase,
                                                        const int* __restrict__ tcount,
                                                        const double* __restrict__ stage, long long nelem,
                                                        double* __restrict__ out0, double* __restrict__ out1,
                                                        double* __restrict__ out2) {
  __shared__ double s1[RED_NE_MAX], s2[RED_NE_MAX];
  const int w0 = blockIdx.x;
  const int wave = UNIFORM ? __builtin_amdgcn_readfirstlane(w0) : w0;
  const int lane = threadIdx.x & 63;
  if (wave >= npair) return;
  const int2 pr = tpairs[wave];
  const bool diag = pr.x == pr.y;
  const Target t1 = tg[pr.x];
  reduce_target(t1, stage + tbase[pr.x], tcount[pr.x], lane, s1);
  Target t2 = t1;
  if (!diag) {
    t2 = tg[pr.y];
    reduce_target(t2, stage + tbase[pr.y], tcount[pr.y], lane, s2);
  }
  __syncthreads();
  // matrix index in pad; element offsets within that matrix
  double* __restrict__ out = t1.pad == 0 ? out0 : (t1.pad == 1 ? out1 : out2);
  const long long o1 = t1.base - (long long)t1.pad * nelem, o2 = t2.base - (long long)t2.pad * nelem;
  const int ne = t1.nrow * t1.ncol;
  for (int e = lane; e < ne; e += 64) {
    const int r = e / t1.ncol, c = e - (e / t1.ncol) * t1.ncol;
    const double v = s1[e] + (diag ? s1[c * t1.ncol + r] : s2[c * t1.nrow + r]);
    out[o1 + (long long)r * t1.stride + c] = v;
    if (!diag) out[o2 + (long long)c * t2.stride + r] = v;
  }
}

hipError_t launch_reduce_sym(int npair, const int2* tpairs, const Target* tg, const long long* tbase,
                             const int* tcount, const double* stage, long long stage_len, long long nelem,
                             double* out0, double* out1, double* out2, hipStream_t s) {
  if (npair <= 0) return hipSuccess;
  // one-wave workgroups: 38 vs 43 us (four waves per group) for the 168 MB
  // staging of a (H2O)16 STO-3G build
  const dim3 grid(npair), block(64);
  // staging larger than the 256 MB Infinity Cache streams from HBM
  if (stage_len * (long long)sizeof(double) > (256LL << 20))
    hipLaunchKernelGGL(reduce_sym_kernel<true>, grid, block, 0, s, npair, tpairs, tg, tbase, tcount, stage, nelem,
                       out0, out1, out2);
  else
    hipLaunchKernelGGL(reduce_sym_kernel<false>, grid, block, 0, s, npair, tpairs, tg, tbase, tcount, stage, nelem,
                       out0, out1, out2);
  return hipGetLastError();
}

// ------------------------------------------------------- Boys hook
template <int L>
__global__ void boys_kernel(const double* __restrict__ T, int n, double* __restrict__ out, int fp32,
                            const double* __restrict__ gtab) {
  __shared__ double btab[BOYS_LDS];
  boys_stage<L>(gtab, btab);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fp32) {
    float F[L + 1];
    boys_fp32<L>((float)T[i], btab, F);
    for (int m = 0; m <= L; ++m) out[(long long)i * (L + 1) + m] = F[m];
  } else {
    double F[L + 1];
    boys_fp64<L>(T[i], btab, F);
    for (int m = 0; m <= L; ++m) out[(long long)i * (L + 1) + m] = F[m];
  }
}

hipError_t launch_boys(const double* T, int n, int mmax, double* out, int fp32, const double* gtab,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  dim3 g((n + 255) / 256), b(256);
  switch (mmax) {
#define B(LL) \
  case LL:    \
    hipLaunchKernelGGL(boys_kernel<LL>, g, b, 0, s, T, n, out, fp32, gtab); break;
    B(0) B(1) B(2) B(3) B(4) B(5) B(6) B(7) B(8) B(9) B(10) B(11) B(12) B(13) B(14) B(15) B(16)
#undef B
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dft
