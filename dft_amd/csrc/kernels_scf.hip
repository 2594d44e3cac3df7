// Symmetric eigensolver for the SCF step: batched two-sided cyclic Jacobi,
// one workgroup per matrix, graph-capturable (no host synchronisation, no
// allocation, a bounded number of sweeps).
//
// It replaces the eigendecompositions of the reference's SCF loops
// (xp.linalg.eigh of X^T F X, shelled.py:1210-1213, main.py:796-799; of S,
// shelled.py:1082 / main.py:754) for matrices up to JAC_NMAX. The matrix
// lives in LDS; the eigenvectors in LDS as well when both fit, else in the
// caller's output buffer (global memory, L2-resident at these sizes).
//
// Method: Golub & Van Loan, Matrix Computations (4th ed.), section 8.5: the
// symmetric 2x2 Schur decomposition (tau = (a_qq - a_pp) / 2 a_pq,
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c)
// applied to m/2 disjoint pairs at once in the round-robin (parallel) order,
// m - 1 rounds per sweep; with J = [[c, s], [-s, c]] on (p, q) the update
// A <- J^T A J touches every 2x2 block {p_k, q_k} x {p_l, q_l} once per
// round. Rotations with |a_pq| below 2^-52 ||A||_F are skipped; the sweeps
// stop after one without any rotation (at most JAC_MAX_SWEEPS). Eigenvalues
// are returned in ascending order, eigenvectors as the columns of V.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dft {

constexpr int JAC_THREADS = 512;
constexpr int JAC_MAX_SWEEPS = 40;
// LDS of one workgroup, and the matrix order up to which A and V share it
constexpr int JAC_LDS_BYTES = 160 * 1024;

// partner positions of round r of the round-robin tournament over m players
__device__ __forceinline__ int rr_pos(int i, int r, int m) { return i == 0 ? 0 : 1 + (i - 1 + r) % (m - 1); }

__global__ __launch_bounds__(JAC_THREADS) void jacobi_eigh_kernel(int n, double* __restrict__ A, long long sa,
                                                                  double* __restrict__ w, long long sw,
                                                                  double* __restrict__ V, long long sv, int init_v,
                                                                  int v_in_lds) {
  extern __shared__ double lds[];
  __shared__ double cs_c[JAC_NMAX / 2], cs_s[JAC_NMAX / 2];
  __shared__ int cs_p[JAC_NMAX / 2], cs_q[JAC_NMAX / 2];
  __shared__ double red[JAC_THREADS / 64];
  __shared__ double diag[JAC_NMAX];
  __shared__ int rotated;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int m = n + (n & 1);  // an odd order gets a decoupled zero row and column
  const int h = m / 2;
  double* a = lds;                                  // m x m
  double* vl = v_in_lds ? lds + (size_t)m * m : nullptr;  // n x n when it fits
  double* Ab = A + (size_t)blockIdx.x * sa;
  double* Vb = V + (size_t)blockIdx.x * sv;
  double* v = v_in_lds ? vl : Vb;
  for (int i = tid; i < m * m; i += nt) {
    const int r = i / m, c = i - r * m;
    a[i] = (r < n && c < n) ? Ab[(size_t)r * n + c] : 0.0;
  }
  for (int i = tid; i < n * n; i += nt) {
    const int r = i / n, c = i - r * n;
    v[i] = init_v ? (r == c ? 1.0 : 0.0) : Vb[i];
  }
  if (tid == 0) rotated = 0;
  __syncthreads();
  // ||A||_F by a fixed-order reduction: per-thread partial, per-wave
  // butterfly, then the waves in order
  double f = 0.0;
  for (int i = tid; i < m * m; i += nt) f = fma(a[i], a[i], f);
  for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = f;
  __syncthreads();
  double fro = 0.0;
  for (int k = 0; k < (nt >> 6); ++k) fro += red[k];
  const double thresh = 2.220446049250313e-16 * sqrt(fro);
  for (int sweep = 0; sweep < JAC_MAX_SWEEPS; ++sweep) {
    for (int r = 0; r < m - 1; ++r) {
      for (int k = tid; k < h; k += nt) {
        int p = rr_pos(k, r, m), q = rr_pos(m - 1 - k, r, m);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        const double apq = a[p * m + q];
        double c = 1.0, s = 0.0;
        if (fabs(apq) > thresh) {
          const double tau = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
          const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
          rotated = 1;
        }
        cs_c[k] = c;
        cs_s[k] = s;
        cs_p[k] = p;
        cs_q[k] = q;
      }
      __syncthreads();
      for (int idx = tid; idx < h * h; idx += nt) {
        const int k = idx / h, l = idx - k * h;
        const double sk = cs_s[k], sl = cs_s[l];
        if (sk == 0.0 && sl == 0.0) continue;
        const double ck = cs_c[k], cl = cs_c[l];
        const int pk = cs_p[k], qk = cs_q[k], pl = cs_p[l], ql = cs_q[l];
        const double xpp = a[pk * m + pl], xpq = a[pk * m + ql], xqp = a[qk * m + pl], xqq = a[qk * m + ql];
        // rows (J^T from the left), then columns (J from the right)
        const double tpp = ck * xpp - sk * xqp, tpq = ck * xpq - sk * xqq;
        const double tqp = sk * xpp + ck * xqp, tqq = sk * xpq + ck * xqq;
        double ypp = cl * tpp - sl * tpq, ypq = sl * tpp + cl * tpq;
        double yqp = cl * tqp - sl * tqq, yqq = sl * tqp + cl * tqq;
        if (k == l) {
          // the block the rotation diagonalises: its off-diagonal is zero
          ypq = 0.0;
          yqp = 0.0;
        }
        a[pk * m + pl] = ypp;
        a[pk * m + ql] = ypq;
        a[qk * m + pl] = yqp;
        a[qk * m + ql] = yqq;
      }
      for (int idx = tid; idx < n * h; idx += nt) {
        const int i = idx / h, k = idx - i * h;
        const double s = cs_s[k];
        const int p = cs_p[k], q = cs_q[k];
        if (s == 0.0 || q >= n) continue;
        const double c = cs_c[k];
        const double vp = v[(size_t)i * n + p], vq = v[(size_t)i * n + q];
        v[(size_t)i * n + p] = c * vp - s * vq;
        v[(size_t)i * n + q] = s * vp + c * vq;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (tid == 0) rotated = 0;
    __syncthreads();
    if (!any) break;
  }
  // ascending order: rank of each eigenvalue (ties by index)
  for (int i = tid; i < n; i += nt) diag[i] = a[i * m + i];
  __syncthreads();
  double* wb = w + (size_t)blockIdx.x * sw;
  if (!v_in_lds) {
    // stage V through the (now free) matrix area to permute its columns
    for (int i = tid; i < n * n; i += nt) a[i] = Vb[i];
    __syncthreads();
    v = a;
  }
  for (int i = tid; i < n; i += nt) {
    const double di = diag[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = diag[j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    wb[rank] = di;
    for (int r = 0; r < n; ++r) Vb[(size_t)r * n + rank] = v[(size_t)r * n + i];
  }
}

hipError_t launch_syevj(int n, int batch, double* A, long long sa, double* w, long long sw, double* V,
                        long long sv, int init_v, hipStream_t s) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (n > JAC_NMAX) return hipErrorInvalidValue;
  const int m = n + (n & 1);
  const size_t a_bytes = (size_t)m * m * sizeof(double), v_bytes = (size_t)n * n * sizeof(double);
  // static LDS of the kernel (rotation tables, diagonal, reduction) stays below 4 KB
  const bool v_in_lds = a_bytes + v_bytes + 4096 <= (size_t)JAC_LDS_BYTES;
  const size_t dyn = v_in_lds ? a_bytes + v_bytes : a_bytes;
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)jacobi_eigh_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       JAC_LDS_BYTES - 4096);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL(jacobi_eigh_kernel, dim3(batch), dim3(JAC_THREADS), dyn, s, n, A, sa, w, sw, V, sv, init_v,
                     v_in_lds ? 1 : 0);
  return hipGetLastError();
}

}  // namespace dft
