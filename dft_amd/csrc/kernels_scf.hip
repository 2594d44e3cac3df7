// Symmetric eigensolver for the SCF step: batched two-sided cyclic Jacobi,
// one workgroup per matrix, graph-capturable (no host synchronisation, no
// allocation, a bounded number of sweeps).
//
// It replaces the eigendecompositions of the reference's SCF loops
// (xp.linalg.eigh of X^T F X, shelled.py:1210-1213, main.py:796-799; of S,
// shelled.py:1082 / main.py:754) for matrices up to JAC_NMAX. The matrix
// lives in LDS; the eigenvectors in LDS as well when both fit, else in the
// caller's output buffer (global memory, L2-resident at these sizes). The
// matrix is stored as its packed upper triangle, so both fit up to n = 114.
//
// Method: Golub & Van Loan, Matrix Computations (4th ed.), section 8.5: the
// symmetric 2x2 Schur decomposition (tau = (a_qq - a_pp) / 2 a_pq,
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c)
// applied to m/2 disjoint pairs at once in the round-robin (parallel) order,
// m - 1 rounds per sweep; with J = [[c, s], [-s, c]] on (p, q) the update
// A <- J^T A J touches every 2x2 block {p_k, q_k} x {p_l, q_l} once per
// round. Rotations with |a_pq| below m 2^-52 ||A||_F are skipped (rounding
// keeps the off-diagonal entries of a converged matrix near 2^-52 ||A|| times
// the square root of the rotations that touched them, so a bound without the
// factor m can be out of reach); the sweeps stop when no off-diagonal entry
// is above that bound, checked in one pass before each sweep (a sweep with
// nothing to rotate is never run), or after a sweep without any rotation
// (at most JAC_MAX_SWEEPS). Eigenvalues
// are returned in ascending order, eigenvectors as the columns of V.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace dft {

constexpr int JAC_THREADS = 1024;
constexpr int JAC_MAX_SWEEPS = 40;
// LDS of one workgroup, and the part of it the kernel's static arrays take
// (rotation tables of two rounds, diagonal, reduction: 5.3 KB)
constexpr int JAC_LDS_BYTES = 160 * 1024;
constexpr int JAC_STATIC_LDS = 6 * 1024;
// per-thread share of a round's blocks (h <= 64 pairs): ceil(64 * 65 / 2 / JAC_THREADS)
constexpr int JAC_BLK_PER_THREAD = (JAC_NMAX / 2 * (JAC_NMAX / 2 + 1) / 2 + JAC_THREADS - 1) / JAC_THREADS;
constexpr int JAC_V_GROUP = 4;

// partner positions of round r < m - 1 of the round-robin tournament over m
// players (i < m, so i - 1 + r < 2 (m - 1) and one subtraction reduces it)
__device__ __forceinline__ int rr_pos(int i, int r, int m) {
  if (i == 0) return 0;
  const int x = i - 1 + r;
  return 1 + (x >= m - 1 ? x - (m - 1) : x);
}

// The pair of round r at table position k, lower index first.
__device__ __forceinline__ int2 rr_pair(int k, int r, int m) {
  const int p = rr_pos(k, r, m), q = rr_pos(m - 1 - k, r, m);
  return p < q ? make_int2(p, q) : make_int2(q, p);
}

// Upper triangle of the symmetric m x m matrix, packed by rows (24-bit
// products: indices and orders stay below 2^12; at lo = 0 the second
// product is 0 whatever lo - 1 wraps to).
__device__ __forceinline__ int tri(int r, int c, int m) {
  const unsigned lo = (unsigned)min(r, c), hi = (unsigned)max(r, c);
  return (int)(__umul24(lo, (unsigned)m) - (__umul24(lo, lo - 1u) >> 1) + (hi - lo));
}

// Rotation log of one matrix (LOG kernels): an int header (the sweeps done,
// then the rank of each column in ascending eigenvalue order), then (c, s)
// of every table position k of every round, JAC_MAX_SWEEPS sweeps of m - 1
// rounds of m / 2 positions; the pair of a position is rr_pair(k, r, m).
constexpr long long JAC_LOG_HEAD = 1024;  // bytes, >= 4 (1 + JAC_NMAX)
long long syevj_log_bytes(int n) {
  const int m = n + (n & 1);
  return JAC_LOG_HEAD + (long long)JAC_MAX_SWEEPS * (m - 1) * (m / 2) * (long long)sizeof(double2);
}

// VL: eigenvectors in LDS (with the packed matrix) or in the output buffer.
// LOG: no eigenvectors here; the rotations go to the log (jlog, jlog_stride
// bytes per matrix) and jacobi_apply_kernel applies them to the rows of V.
template <bool VL, bool LOG>
__global__ __launch_bounds__(JAC_THREADS) void jacobi_eigh_kernel(int n, const double* __restrict__ A, long long sa,
                                                                  double* __restrict__ w, long long sw,
                                                                  double* __restrict__ V, long long sv,
                                                                  int init_v, int* __restrict__ sweeps, double tol,
                                                                  char* __restrict__ jlog, long long jlog_stride) {
  extern __shared__ double lds[];
  __shared__ double cs_c[2][JAC_NMAX / 2], cs_s[2][JAC_NMAX / 2], cs_t[2][JAC_NMAX / 2];
  __shared__ int cs_p[2][JAC_NMAX / 2], cs_q[2][JAC_NMAX / 2];
  __shared__ double red[JAC_THREADS / 64];
  __shared__ double diag[JAC_NMAX];
  __shared__ int rotated;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int m = n + (n & 1);  // an odd order gets a decoupled zero row and column
  const int h = m / 2;
  const int ntri = m * (m + 1) / 2;
  double* a = lds;  // packed upper triangle, ntri
  const double* Ab = A + (size_t)blockIdx.x * sa;
  double* __restrict__ Vb = V + (size_t)blockIdx.x * sv;
  double* vl = lds + ntri;  // n x n (VL only)
  char* lg = LOG ? jlog + (size_t)blockIdx.x * jlog_stride : nullptr;
  double2* lcs = LOG ? (double2*)(lg + JAC_LOG_HEAD) : nullptr;
  for (int i = tid; i < m * m; i += nt) {
    const int r = i / m, c = i - r * m;
    if (r <= c) a[tri(r, c, m)] = (r < n && c < n) ? Ab[(size_t)r * n + c] : 0.0;
  }
  if (!LOG)
    for (int i = tid; i < n * n; i += nt) {
      const int r = i / n, c = i - r * n;
      const double x = init_v ? (r == c ? 1.0 : 0.0) : Vb[i];
      if (VL) vl[i] = x;
      else Vb[i] = x;
    }
  if (tid == 0) rotated = 0;
  __syncthreads();
  // ||A||_F by a fixed-order reduction: per-thread partial, per-wave
  // butterfly, then the waves in order (off-diagonal entries count twice)
  double f = 0.0;
  for (int i = tid; i < m * m; i += nt) {
    const int r = i / m, c = i - r * m;
    const double x = a[tri(r, c, m)];
    f = fma(x, x, f);
  }
  for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = f;
  __syncthreads();
  double fro = 0.0;
  for (int k = 0; k < (nt >> 6); ++k) fro += red[k];
  // the rotation / convergence bound: m 2^-52 ||A||_F, or the caller's
  // absolute tolerance on the off-diagonal entries where that is larger
  const double thresh = fmax(m * 2.220446049250313e-16 * sqrt(fro), tol);
  // this thread's fixed share of the round's work: blocks (k, l), k <= l, of
  // the h x h block grid, and (row, pair) updates of the eigenvectors
  const int nblk = h * (h + 1) / 2;
  // (packed: block k | l << 8)
  int bkl[JAC_BLK_PER_THREAD], nb = 0;
  for (int idx = tid; idx < nblk; idx += nt) {
    int k = 0, rem = idx;
    while (rem >= h - k) {
      rem -= h - k;
      ++k;
    }
    bkl[nb++] = k | (k + rem) << 8;
  }
  // eigenvector update by the threads of waves 1.. (wave 0 computes the
  // next round's rotations meanwhile): vng row groups per pair
  const int vu = tid - 64, vng = (nt - 64) / h;
  const int vk = (vu >= 0 && vu < vng * h) ? vu % h : -1, vg = vu >= 0 ? vu / h : 0;
  int done = 0;  // sweeps performed
  for (int sweep = 0; sweep < JAC_MAX_SWEEPS; ++sweep) {
    // a sweep that would rotate no pair (every |a_pq| <= thresh) changes
    // nothing: stop before it, from one pass over the off-diagonal entries
    {
      double om = 0.0;
      for (int i = tid; i < m * m; i += nt) {
        const int r = i / m, c = i - r * m;
        if (r < c) om = fmax(om, fabs(a[tri(r, c, m)]));
      }
      for (int off = 32; off > 0; off >>= 1) om = fmax(om, __shfl_xor(om, off, 64));
      __syncthreads();
      if ((tid & 63) == 0) red[tid >> 6] = om;
      __syncthreads();
      double mx = 0.0;
      for (int k = 0; k < (nt >> 6); ++k) mx = fmax(mx, red[k]);
      if (!(mx > thresh)) break;
    }
    // Round r: wave 0 computes the rotations of round r (disjoint pairs)
    // while the other waves apply round r - 1's rotations to the
    // eigenvector columns (double-buffered rotation tables); then every
    // thread updates its blocks of A with round r's rotations. Pass m - 1
    // only finishes the eigenvectors of the sweep's last round.
    for (int r = 0; r < m; ++r) {
      const int cur = r & 1, prev = cur ^ 1;
      if (tid < 64) {
        if (r < m - 1)
          for (int k = tid; k < h; k += 64) {
            const int2 pq = rr_pair(k, r, m);
            const int p = pq.x, q = pq.y;
            const double apq = a[tri(p, q, m)];
            double c = 1.0, s = 0.0, t = 0.0;
            if (fabs(apq) > thresh) {
              // t = sign(tau) / (|tau| + sqrt(1 + tau^2)) with tau = d / e,
              // d = a_qq - a_pp, e = 2 a_pq, written as sign(d) e / (|d| + |(d, e)|)
              const double d = a[tri(q, q, m)] - a[tri(p, p, m)], e = 2.0 * apq;
              t = (d >= 0.0 ? e : -e) / (fabs(d) + sqrt(fma(d, d, e * e)));
              c = 1.0 / sqrt(fma(t, t, 1.0));
              s = t * c;
              rotated = 1;
            }
            cs_c[cur][k] = c;
            cs_s[cur][k] = s;
            cs_t[cur][k] = t;
            cs_p[cur][k] = p;
            cs_q[cur][k] = q;
            // logged for the order of the tournament slots (k, m - 1 - k),
            // which jacobi_apply_kernel keeps in one lane: s, or -s where
            // slot k holds the larger index
            if (LOG)
              lcs[(size_t)(sweep * (m - 1) + r) * h + k] =
                  make_double2(c, rr_pos(k, r, m) < rr_pos(m - 1 - k, r, m) ? s : -s);
          }
      } else if (!LOG && r > 0 && vk >= 0) {
        // eigenvector columns p_k, q_k of rows g, g + ng, ... for this
        // thread's fixed pair k and row group g (its rotation read once), in
        // groups of JAC_V_GROUP rows loaded before any is stored
        const double sr = cs_s[prev][vk];
        const int p = cs_p[prev][vk], q = cs_q[prev][vk];
        if (sr != 0.0 && q < n) {
          const double c = cs_c[prev][vk];
          for (int i0 = vg; i0 < n; i0 += JAC_V_GROUP * vng) {
            double xp[JAC_V_GROUP], xq[JAC_V_GROUP];
#pragma unroll
            for (int u = 0; u < JAC_V_GROUP; ++u) {
              const int i = i0 + u * vng;
              if (i < n) {
                xp[u] = VL ? vl[i * n + p] : Vb[(size_t)i * n + p];
                xq[u] = VL ? vl[i * n + q] : Vb[(size_t)i * n + q];
              }
            }
#pragma unroll
            for (int u = 0; u < JAC_V_GROUP; ++u) {
              const int i = i0 + u * vng;
              if (i < n) {
                // explicit fused forms, shared with jacobi_apply_kernel (c
                // multiplies the entry itself in both, so swapping p and q
                // is the same arithmetic with -s)
                const double np_ = fma(c, xp[u], -(sr * xq[u])), nq_ = fma(c, xq[u], sr * xp[u]);
                if (VL) {
                  vl[i * n + p] = np_;
                  vl[i * n + q] = nq_;
                } else {
                  Vb[(size_t)i * n + p] = np_;
                  Vb[(size_t)i * n + q] = nq_;
                }
              }
            }
          }
        }
      }
      __syncthreads();
      if (r == m - 1) break;
      // blocks {p_k, q_k} x {p_l, q_l} with k <= l (the others are their
      // transposes). The entries a thread touches in a round are its own, so
      // all of them are loaded first and stored after the arithmetic (the
      // compiler cannot see that the stores do not alias later loads).
      {
        double x[JAC_BLK_PER_THREAD][4];
        int ix[JAC_BLK_PER_THREAD][4];
#pragma unroll
        for (int j = 0; j < JAC_BLK_PER_THREAD; ++j) {
          if (j < nb) {
            const int k = bkl[j] & 255, l = bkl[j] >> 8;
            // the pairs of this round, computed here rather than read from
            // the rotation table
            const int2 pqk = rr_pair(k, r, m), pql = rr_pair(l, r, m);
            const int pk = pqk.x, qk = pqk.y, pl = pql.x, ql = pql.y;
            if (k == l) {
              ix[j][0] = tri(pk, pk, m);
              ix[j][1] = tri(pk, qk, m);
              ix[j][2] = tri(qk, qk, m);
              ix[j][3] = ix[j][1];
            } else {
              ix[j][0] = tri(pk, pl, m);
              ix[j][1] = tri(pk, ql, m);
              ix[j][2] = tri(qk, pl, m);
              ix[j][3] = tri(qk, ql, m);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) x[j][e] = a[ix[j][e]];
          }
        }
#pragma unroll
        for (int j = 0; j < JAC_BLK_PER_THREAD; ++j) {
          if (j < nb) {
            const int k = bkl[j] & 255, l = bkl[j] >> 8;
            const double sk = cs_s[cur][k], sl = cs_s[cur][l];
            if (sk == 0.0 && sl == 0.0) continue;
            const double ck = cs_c[cur][k], cl = cs_c[cur][l];
            if (k == l) {
              // the block the rotation diagonalises: a_pp - t a_pq, a_qq + t a_pq, 0
              const double t = cs_t[cur][k], apq = x[j][1];
              a[ix[j][0]] = x[j][0] - t * apq;
              a[ix[j][2]] = x[j][2] + t * apq;
              a[ix[j][1]] = 0.0;
              continue;
            }
            // rows (J^T from the left), then columns (J from the right)
            const double tpp = ck * x[j][0] - sk * x[j][2], tpq = ck * x[j][1] - sk * x[j][3];
            const double tqp = sk * x[j][0] + ck * x[j][2], tqq = sk * x[j][1] + ck * x[j][3];
            a[ix[j][0]] = cl * tpp - sl * tpq;
            a[ix[j][1]] = sl * tpp + cl * tpq;
            a[ix[j][2]] = cl * tqp - sl * tqq;
            a[ix[j][3]] = sl * tqp + cl * tqq;
          }
        }
      }
      __syncthreads();
    }
    ++done;
    const int any = rotated;
    __syncthreads();
    if (tid == 0) rotated = 0;
    __syncthreads();
    if (!any) break;
  }
  if (sweeps && tid == 0) sweeps[blockIdx.x] = done;
  if (LOG && tid == 0) ((int*)lg)[0] = done;
  // ascending order: rank of each eigenvalue (ties by index)
  for (int i = tid; i < n; i += nt) diag[i] = a[tri(i, i, m)];
  __syncthreads();
  double* wb = w + (size_t)blockIdx.x * sw;
  if (LOG) {
    for (int i = tid; i < n; i += nt) {
      const double di = diag[i];
      int rank = 0;
      for (int j = 0; j < n; ++j) {
        const double dj = diag[j];
        rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
      }
      wb[rank] = di;
      ((int*)lg)[1 + i] = rank;
    }
    return;
  }
  const double* v = vl;
  if (!VL) {
    // stage V through LDS (the packed matrix is no longer needed) to permute its columns
    double* st = lds;
    for (int i = tid; i < n * n; i += nt) st[i] = Vb[i];
    __syncthreads();
    v = st;
  }
  for (int i = tid; i < n; i += nt) {
    const double di = diag[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = diag[j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    wb[rank] = di;
    for (int r = 0; r < n; ++r) Vb[(size_t)r * n + rank] = v[r * n + i];
  }
}

// A double moved one lane along the wave by DPP (wave_shl1 = 0x130: lane k
// takes lane k + 1's value; wave_shr1 = 0x138: lane k takes lane k - 1's;
// a lane without a source gets 0).
template <int CTRL>
__device__ __forceinline__ double wave_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The eigenvectors from a rotation log: one wave per row of V, the row in
// registers. Lane k holds the entries at tournament slots k and m - 1 - k,
// which form pair k of every round; slot 0 keeps index 0 and the indices at
// slots 1 .. m - 1 move one slot down per round (rr_pos(i, r + 1) =
// rr_pos(i + 1, r), cyclically), so between rounds the first entries shift
// one lane down and the second entries one lane up (DPP), with the two ends
// of the ring turning round inside lanes h - 1 and 0. The log holds s signed
// for the slot order, so a round is the in-workgroup update's explicit fused
// forms with no selects (the same V bit for bit, up to the sign of a zero
// where a pair is skipped). A pair with the padding index n of an odd order
// is never rotated (its entries are 0), and no lane below h reads a lane at
// or above h. After the logged whole sweeps every slot holds its own index
// again, and the row is written with its columns in ascending eigenvalue
// order.
constexpr int JAC_APPLY_ROWS = 4;    // rows (waves) per workgroup
constexpr int JAC_APPLY_AHEAD = 8;   // rounds of (c, s) loaded ahead

__device__ __forceinline__ void jac_apply_round(double& xa, double& xb, double2 g, int lane, int h) {
  const double na = fma(g.x, xa, -(g.y * xb)), nb = fma(g.x, xb, g.y * xa);
  const double df = wave_shift<0x130>(na), us = wave_shift<0x138>(nb);
  // h = 1 (orders 1 and 2): a ring of one slot, nothing moves
  xa = lane == 0 ? na : (lane == h - 1 ? nb : df);
  xb = lane == 0 ? (h > 1 ? df : nb) : us;
}

__global__ __launch_bounds__(64 * JAC_APPLY_ROWS) void jacobi_apply_kernel(int n, const char* __restrict__ jlog,
                                                                           long long jlog_stride,
                                                                           double* __restrict__ V, long long sv,
                                                                           int init_v) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * JAC_APPLY_ROWS + wv;
  if (i >= n) return;  // whole waves: the shifts below run with all 64 lanes
  const char* lg = jlog + (size_t)blockIdx.y * jlog_stride;
  const int* head = (const int*)lg;
  const double2* lcs = (const double2*)(lg + JAC_LOG_HEAD);
  double* __restrict__ Vr = V + (size_t)blockIdx.y * sv + (size_t)i * n;
  const int m = n + (n & 1), h = m / 2;
  const bool act = lane < h;
  // round 0: slot k holds index k (k < h <= n); index m - 1 = n of an odd
  // order is the decoupled padding entry, 0 throughout
  const int ib = m - 1 - lane;
  double xa = 0.0, xb = 0.0;
  if (act) {
    xa = init_v ? (i == lane ? 1.0 : 0.0) : Vr[lane];
    if (ib < n) xb = init_v ? (i == ib ? 1.0 : 0.0) : Vr[ib];
  }
  const int nround = head[0] * (m - 1);
  // loads stay inside the log (clamped round and position; what a load past
  // the logged rounds or of an idle lane returns is never used)
  const int last = JAC_MAX_SWEEPS * (m - 1) - 1, kk = min(lane, h - 1);
  auto load = [&](int t) { return lcs[(size_t)min(t, last) * h + kk]; };
  double2 g[JAC_APPLY_AHEAD], gn[JAC_APPLY_AHEAD];
#pragma unroll
  for (int u = 0; u < JAC_APPLY_AHEAD; ++u) g[u] = load(u);
  const int nfull = nround - nround % JAC_APPLY_AHEAD;
  int t0 = 0;
  for (; t0 < nfull; t0 += JAC_APPLY_AHEAD) {
#pragma unroll
    for (int u = 0; u < JAC_APPLY_AHEAD; ++u) gn[u] = load(t0 + JAC_APPLY_AHEAD + u);
#pragma unroll
    for (int u = 0; u < JAC_APPLY_AHEAD; ++u) jac_apply_round(xa, xb, g[u], lane, h);
#pragma unroll
    for (int u = 0; u < JAC_APPLY_AHEAD; ++u) g[u] = gn[u];
  }
  for (int t = t0; t < nround; ++t) jac_apply_round(xa, xb, load(t), lane, h);
  if (act) {
    Vr[head[1 + lane]] = xa;
    if (ib < n) Vr[head[1 + ib]] = xb;
  }
}

hipError_t launch_syevj(int n, int batch, double* A, long long sa, double* w, long long sw, double* V,
                        long long sv, int init_v, int* sweeps, double tol, char* jlog, long long jlog_stride,
                        hipStream_t s) {
  if (batch <= 0 || n <= 0) return hipSuccess;
  if (n > JAC_NMAX) return hipErrorInvalidValue;
  if (jlog && jlog_stride < syevj_log_bytes(n)) return hipErrorInvalidValue;
  const int m = n + (n & 1);
  const size_t a_bytes = (size_t)m * (m + 1) / 2 * sizeof(double), v_bytes = (size_t)n * n * sizeof(double);
  const bool vl = a_bytes + v_bytes + JAC_STATIC_LDS <= (size_t)JAC_LDS_BYTES;
  // the staging of V for the final column permutation reuses the matrix area
  const size_t dyn = jlog ? a_bytes : vl ? a_bytes + v_bytes : std::max(a_bytes, v_bytes);
  static bool attr_set = false;
  if (!attr_set) {
    for (const void* k : {(const void*)jacobi_eigh_kernel<true, false>, (const void*)jacobi_eigh_kernel<false, false>,
                          (const void*)jacobi_eigh_kernel<false, true>}) {
      hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, JAC_LDS_BYTES - JAC_STATIC_LDS);
      if (e != hipSuccess) return e;
    }
    attr_set = true;
  }
  if (jlog) {
    hipLaunchKernelGGL((jacobi_eigh_kernel<false, true>), dim3(batch), dim3(JAC_THREADS), dyn, s, n, A, sa, w, sw, V,
                       sv, init_v, sweeps, tol, jlog, jlog_stride);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(jacobi_apply_kernel, dim3((n + JAC_APPLY_ROWS - 1) / JAC_APPLY_ROWS, batch),
                       dim3(64 * JAC_APPLY_ROWS), 0, s, n, (const char*)jlog, jlog_stride, V, sv, init_v);
  } else if (vl) {
    hipLaunchKernelGGL((jacobi_eigh_kernel<true, false>), dim3(batch), dim3(JAC_THREADS), dyn, s, n, A, sa, w, sw, V,
                       sv, init_v, sweeps, tol, (char*)nullptr, 0LL);
  } else {
    hipLaunchKernelGGL((jacobi_eigh_kernel<false, false>), dim3(batch), dim3(JAC_THREADS), dyn, s, n, A, sa, w, sw,
                       V, sv, init_v, sweeps, tol, (char*)nullptr, 0LL);
  }
  return hipGetLastError();
}

}  // namespace dft

namespace dft {

// ---------------------------------------------------------------------------
// One damped + DIIS SCF iteration on the device (closed shell), as the loop of
// shelled.py:1164-1260 with P_a = P_b: the two kernels below hold the
// loop's control state in device memory (st[], SCF_ST_* slots), so an
// iteration needs no host decision and can be replayed as a hipGraph; the
// host reads st[] once per iteration to stop the loop.

constexpr int SCF_THREADS = 1024;

__device__ __forceinline__ double block_sum(double v, double* red) {
  // fixed order: wave butterflies, then the waves in order
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
  return s;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s = fmax(s, red[k]);
  return s;
}

// History (ring of h slots, this iteration's slot = count mod h): store F and
// its residual R = F P S - S P F, update the Gram matrix G[s][t] = <R_s, R_t>
// of the new slot's row and column, solve the DIIS equations on the used
// slots (B c = 1 without a Lagrange row, c normalised to sum 1; unused slots
// get c = 0), and form the Fock matrix of this iteration Fu (the
// extrapolation once DIIS is on, F before), and E_sum = sum P.H + 2 sum Pa.Fu.
// Four launches: the element-wise passes (history store + Gram row partials,
// extrapolation + energy partials) spread over scf_nblk(n) workgroups, each
// writing its partial sums to its own state slots; the partials are summed
// in block order by single-workgroup kernels (results independent of timing).
constexpr int SCF_EW_THREADS = 256;
constexpr int SCF_NBLK_MAX = 64;
constexpr int SCF_PART_STRIDE = SCF_HMAX + 1;  // per block: h Gram partials, then the energy partial
static_assert(SCF_ST_PART + SCF_NBLK_MAX * SCF_PART_STRIDE <= SCF_STATE_LEN, "SCF state too short");

__host__ __device__ inline int scf_nblk(int n) {
  const long long nn = (long long)n * n;
  const long long b = (nn + 4 * SCF_EW_THREADS - 1) / (4 * SCF_EW_THREADS);
  return (int)(b < 1 ? 1 : (b > SCF_NBLK_MAX ? SCF_NBLK_MAX : b));
}

// fixed-order sum over the block: wave butterflies, then the waves in order
__device__ __forceinline__ double ew_block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
  return s;
}

// elements [b * chunk, (b + 1) * chunk) of block b
__device__ __forceinline__ void ew_range(long long nn, long long& i0, long long& i1) {
  const long long chunk = (nn + gridDim.x - 1) / gridDim.x;
  i0 = (long long)blockIdx.x * chunk;
  i1 = i0 + chunk < nn ? i0 + chunk : nn;
}

__global__ __launch_bounds__(SCF_EW_THREADS) void scf_diis_dots_kernel(int n, int h, const double* __restrict__ F,
                                                                       const double* __restrict__ R,
                                                                       double* __restrict__ histF,
                                                                       double* __restrict__ histR,
                                                                       double* __restrict__ st) {
  const long long nn = (long long)n * n;
  const int slot = (int)st[SCF_ST_COUNT] % h;
  long long i0, i1;
  ew_range(nn, i0, i1);
  double d[SCF_HMAX];
#pragma unroll
  for (int t = 0; t < SCF_HMAX; ++t) d[t] = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double r = R[i];
    histF[slot * nn + i] = F[i];
    histR[slot * nn + i] = r;
#pragma unroll
    for (int t = 0; t < SCF_HMAX; ++t)
      if (t < h) d[t] = fma(r, t == slot ? r : histR[t * nn + i], d[t]);
  }
  // the block's h dot products together: each wave's butterflies, one
  // barrier, then the waves in order (the additions of ew_block_sum, which
  // would take two barriers per product)
  __shared__ double redv[SCF_EW_THREADS / 64][SCF_HMAX];
#pragma unroll
  for (int t = 0; t < SCF_HMAX; ++t)
    if (t < h) {
      double v = d[t];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((threadIdx.x & 63) == 0) redv[threadIdx.x >> 6][t] = v;
    }
  __syncthreads();
  double* part = st + SCF_ST_PART + blockIdx.x * SCF_PART_STRIDE;
  if ((int)threadIdx.x < h) {
    double x = 0.0;
    for (int k = 0; k < SCF_EW_THREADS / 64; ++k) x += redv[k][threadIdx.x];
    part[threadIdx.x] = x;
  }
}

// sum_b p[b * SCF_PART_STRIDE] over b < nblk, in b order, with the loads
// issued 16 at a time ahead of the adds (a loop that adds each term as it
// arrives waits one memory latency per term: 64 terms, about 20 us)
__device__ __forceinline__ double scf_part_sum(const double* __restrict__ p, int nblk) {
  constexpr int B = 16;
  double s = 0.0;
  for (int b0 = 0; b0 < nblk; b0 += B) {
    double v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) v[u] = b0 + u < nblk ? p[(size_t)(b0 + u) * SCF_PART_STRIDE] : 0.0;
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (b0 + u < nblk) s += v[u];
  }
  return s;
}

__global__ __launch_bounds__(64) void scf_diis_solve_kernel(int h, int nblk, double* __restrict__ G,
                                                            double* __restrict__ st) {
  const int tid = threadIdx.x;
  const int count = (int)st[SCF_ST_COUNT];
  const int slot = count % h;
  const int used = min(count + 1, h);
  if (tid < h) {
    const double x = scf_part_sum(st + SCF_ST_PART + tid, nblk);
    G[slot * h + tid] = x;
    G[tid * h + slot] = x;
  }
  __syncthreads();
  // Gaussian elimination with partial pivoting on the padded system (rows
  // beyond `used` are identity rows with a zero right-hand side), one row
  // per lane in registers: the pivot search is a wave reduction (largest
  // |M[i][k]|, the lowest row on ties), a row swap and the pivot row's
  // broadcast are lane shuffles, and every lane updates its own row with
  // the same operations, in the same order, as a sequential elimination.
  const int lane = tid;
  double row[SCF_HMAX + 1];
#pragma unroll
  for (int j = 0; j <= SCF_HMAX; ++j) {
    double x = 0.0;
    if (lane < h && j < h) x = (lane < used && j < used) ? G[lane * h + j] : (lane == j ? 1.0 : 0.0);
    if (lane < h && j == h) x = lane < used ? 1.0 : 0.0;
    row[j] = x;
  }
  for (int k = 0; k < h; ++k) {
    // pivot: largest |M[i][k]| over rows i >= k (first such row)
    double mag = (lane >= k && lane < h) ? fabs(row[0]) : -1.0;
#pragma unroll
    for (int j = 1; j <= SCF_HMAX; ++j)
      if (j == k) mag = (lane >= k && lane < h) ? fabs(row[j]) : -1.0;
    int piv = lane;
    for (int off = 32; off > 0; off >>= 1) {
      const double om = __shfl_xor(mag, off, 64);
      const int op = __shfl_xor(piv, off, 64);
      if (om > mag || (om == mag && op < piv)) {
        mag = om;
        piv = op;
      }
    }
    // rows k and piv trade places
    const int src = lane == k ? piv : (lane == piv ? k : lane);
#pragma unroll
    for (int j = 0; j <= SCF_HMAX; ++j) row[j] = __shfl(row[j], src, 64);
    // pivot row broadcast, then rows below it eliminate column k
    double pr[SCF_HMAX + 1];
#pragma unroll
    for (int j = 0; j <= SCF_HMAX; ++j) pr[j] = __shfl(row[j], k, 64);
    double d = pr[0], mk = row[0];
#pragma unroll
    for (int j = 1; j <= SCF_HMAX; ++j)
      if (j == k) {
        d = pr[j];
        mk = row[j];
      }
    if (lane > k && lane < h) {
      const double f = mk / d;
#pragma unroll
      for (int j = 0; j <= SCF_HMAX; ++j)
        if (j >= k && j <= h) row[j] -= f * pr[j];
    }
  }
  // back substitution: c[k] from row k (lane k), broadcast to every lane
  double c[SCF_HMAX];
#pragma unroll
  for (int j = 0; j < SCF_HMAX; ++j) c[j] = 0.0;
  for (int k = h - 1; k >= 0; --k) {
    double x = 0.0, diag = 1.0;
#pragma unroll
    for (int j = 0; j <= SCF_HMAX; ++j)
      if (j == h) x = row[j];
#pragma unroll
    for (int j = 0; j < SCF_HMAX; ++j)
      if (j > k && j < h) x -= row[j] * c[j];
#pragma unroll
    for (int j = 0; j < SCF_HMAX; ++j)
      if (j == k) diag = row[j];
    const double ck = __shfl(x / diag, k, 64);
#pragma unroll
    for (int j = 0; j < SCF_HMAX; ++j)
      if (j == k) c[j] = ck;
  }
  if (lane == 0) {
    double* co = st + SCF_ST_COEF;
    double sum = 0.0;
    for (int k = 0; k < h; ++k) sum += c[k];
    for (int k = 0; k < h; ++k) co[k] = c[k] / sum;
    // A singular Gram matrix (residuals that no longer differ, as near a
    // tight convergence) has no DIIS solution: the iteration then uses its
    // own Fock matrix (c = e_slot) instead of non-finite coefficients.
    bool ok = isfinite(sum) && sum != 0.0;
    for (int k = 0; k < h; ++k) ok = ok && isfinite(co[k]);
    if (!ok)
      for (int k = 0; k < h; ++k) co[k] = k == slot ? 1.0 : 0.0;
    st[SCF_ST_SLOT] = slot;
  }
}

__global__ __launch_bounds__(SCF_EW_THREADS) void scf_diis_extrap_kernel(int n, int h, const double* __restrict__ F,
                                                                         const double* __restrict__ histF,
                                                                         const double* __restrict__ P,
                                                                         const double* __restrict__ H,
                                                                         const double* __restrict__ Pa,
                                                                         double* __restrict__ st,
                                                                         double* __restrict__ Fu) {
  __shared__ double red[SCF_EW_THREADS / 64];
  __shared__ double c[SCF_HMAX];
  const long long nn = (long long)n * n;
  const int used = min((int)st[SCF_ST_COUNT] + 1, h);
  const bool diis = st[SCF_ST_USE_DIIS] != 0.0;
  if (threadIdx.x < SCF_HMAX) c[threadIdx.x] = threadIdx.x < h ? st[SCF_ST_COEF + threadIdx.x] : 0.0;
  __syncthreads();
  long long i0, i1;
  ew_range(nn, i0, i1);
  double e = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double x = F[i];
    if (diis) {
      x = 0.0;
      for (int t = 0; t < used; ++t) x = fma(c[t], histF[t * nn + i], x);
    }
    Fu[i] = x;
    e = fma(P[i], H[i], e) + 2.0 * Pa[i] * x;
  }
  e = ew_block_sum(e, red);
  if (threadIdx.x == 0) st[SCF_ST_PART + blockIdx.x * SCF_PART_STRIDE + SCF_HMAX] = e;
}

__global__ void scf_energy_sum_kernel(int nblk, double* __restrict__ st) {
  if (threadIdx.x != 0) return;
  st[SCF_ST_ESUM] = scf_part_sum(st + SCF_ST_PART + SCF_HMAX, nblk);
}

// Energy and convergence of the iteration, the damped density
// Pa' = rate Pauf + (1 - rate) Pa with dP = max |Pa' - Pa|, the strike and
// DIIS-switch rules (shelled.py:1250-1257) and the next iteration's
// densities (Pa <- Pa', P <- 2 Pa'), unless this iteration converged: then
// Pa keeps the density the final rebuild uses (shelled.py:1261-1285).
__global__ __launch_bounds__(SCF_THREADS) void scf_damp_control_kernel(int n, double* __restrict__ Pa,
                                                                       const double* __restrict__ Pauf,
                                                                       double* __restrict__ P,
                                                                       double* __restrict__ st) {
  __shared__ double red[SCF_THREADS / 64];
  __shared__ int conv;
  // an iteration after the converged one (a host loop that queues the next
  // iteration before it reads this one's state) leaves the loop state and
  // the densities as they are
  if (st[SCF_ST_CONV] != 0.0) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const long long nn = (long long)n * n;
  const double rate = st[SCF_ST_RATE];
  double d = 0.0;
  for (long long i = tid; i < nn; i += nt) {
    const double x = rate * Pauf[i] + (1.0 - rate) * Pa[i];
    d = fmax(d, fabs(x - Pa[i]));
  }
  const double dP = block_max(d, red);
  if (tid == 0) {
    const double E = 0.5 * st[SCF_ST_ESUM] + st[SCF_ST_ENN];
    const double dE = st[SCF_ST_HAS_E] != 0.0 ? fabs(E - st[SCF_ST_E]) : __builtin_inf();
    const bool diis = st[SCF_ST_USE_DIIS] != 0.0;
    if ((dE > st[SCF_ST_PREV_DE] || dP > st[SCF_ST_PREV_DP]) && !diis) {
      if (st[SCF_ST_STRIKES] == 3.0) {
        st[SCF_ST_RATE] = rate / 1.1;
        st[SCF_ST_STRIKES] = 0.0;
      } else {
        st[SCF_ST_STRIKES] += 1.0;
      }
    }
    if (dE < st[SCF_ST_DIIS_TH]) st[SCF_ST_USE_DIIS] = 1.0;
    const bool c = dE < st[SCF_ST_ETOL] && dP < st[SCF_ST_PTOL];
    st[SCF_ST_ENEW] = E;
    st[SCF_ST_DE] = dE;
    st[SCF_ST_DP] = dP;
    st[SCF_ST_CONV] = c ? 1.0 : 0.0;
    if (!c) {
      st[SCF_ST_E] = E;
      st[SCF_ST_HAS_E] = 1.0;
      st[SCF_ST_PREV_DE] = dE;
      st[SCF_ST_PREV_DP] = dP;
      st[SCF_ST_COUNT] += 1.0;
    }
    conv = c ? 1 : 0;
  }
  __syncthreads();
  if (conv) return;
  for (long long i = tid; i < nn; i += nt) {
    const double x = rate * Pauf[i] + (1.0 - rate) * Pa[i];
    Pa[i] = x;
    P[i] = 2.0 * x;
  }
}

hipError_t launch_scf_diis_stage(int n, int h, const double* F, const double* R, double* histF, double* histR,
                                 double* G, const double* P, const double* H, const double* Pa, double* st,
                                 double* Fu, hipStream_t s) {
  if (h < 1 || h > SCF_HMAX || n < 1) return hipErrorInvalidValue;
  const int nb = scf_nblk(n);
  hipLaunchKernelGGL(scf_diis_dots_kernel, dim3(nb), dim3(SCF_EW_THREADS), 0, s, n, h, F, R, histF, histR, st);
  hipLaunchKernelGGL(scf_diis_solve_kernel, dim3(1), dim3(64), 0, s, h, nb, G, st);
  hipLaunchKernelGGL(scf_diis_extrap_kernel, dim3(nb), dim3(SCF_EW_THREADS), 0, s, n, h, F, histF, P, H, Pa, st, Fu);
  hipLaunchKernelGGL(scf_energy_sum_kernel, dim3(1), dim3(64), 0, s, nb, st);
  return hipGetLastError();
}

hipError_t launch_scf_damp_control(int n, double* Pa, const double* Pauf, double* P, double* st, hipStream_t s) {
  hipLaunchKernelGGL(scf_damp_control_kernel, dim3(1), dim3(SCF_THREADS), 0, s, n, Pa, Pauf, P, st);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Dense products of the SCF iteration (orders up to JAC_NMAX): one 256-thread
// workgroup per 16 x 16 output tile ((n / 16)^2 workgroups: 49 at n = 112, a
// library GEMM's 128 x 64 tiles would use 2), the operand panels staged in
// LDS as [k][16] columns, each output one thread's FMA chain over k in order.

constexpr int SG_T = 16;
constexpr int SG_THREADS = SG_T * SG_T;
typedef double SgPanel[JAC_NMAX][SG_T + 1];

// P[k][c] = M(k, idx0 + c) (col = true: column idx0 + c of row-major M) or
// M(idx0 + c, k) (col = false: row idx0 + c), k < kd; entries past nlim
// read as 0. Consecutive threads walk the contiguous index.
__device__ __forceinline__ void sg_panel(SgPanel& P, const double* __restrict__ M, int ld, bool col, int idx0,
                                         int kd, int nlim) {
  for (int e = threadIdx.x; e < kd * SG_T; e += SG_THREADS) {
    int k, c;
    if (col) {
      k = e / SG_T;
      c = e - k * SG_T;
    } else {
      c = e / kd;
      k = e - c * kd;
    }
    const int x = idx0 + c;
    P[k][c] = x < nlim ? (col ? M[(size_t)k * ld + x] : M[(size_t)x * ld + k]) : 0.0;
  }
}

// The same with M = (H + J) - K formed on the fly (the Fock matrix, added
// in that order).
__device__ __forceinline__ void sg_panel_fock(SgPanel& P, const double* __restrict__ H, const double* __restrict__ J,
                                              const double* __restrict__ K, int n, bool col, int idx0) {
  for (int e = threadIdx.x; e < n * SG_T; e += SG_THREADS) {
    int k, c;
    if (col) {
      k = e / SG_T;
      c = e - k * SG_T;
    } else {
      c = e / n;
      k = e - c * n;
    }
    const int x = idx0 + c;
    double v = 0.0;
    if (x < n) {
      const size_t a = col ? (size_t)k * n + x : (size_t)x * n + k;
      v = (H[a] + J[a]) - K[a];
    }
    P[k][c] = v;
  }
}

// sum_k A[k][r] B[k][c] in k order
__device__ __forceinline__ double sg_dot(const SgPanel& A, const SgPanel& B, int r, int c, int kd) {
  double s = 0.0;
  for (int k = 0; k < kd; ++k) s = fma(A[k][r], B[k][c], s);
  return s;
}

// C (m x nc, ldc) = op(A) op(B) over kd; op(A) = A (m x kd, lda) or A^T
// (A kd x m); op(B) = B (kd x nc, ldb) or B^T (B nc x kd). With cp_src, the
// workgroup also copies its tile of cp_src to cp_dst (same shape and ldc).
__global__ __launch_bounds__(SG_THREADS) void scf_gemm_kernel(int m, int nc, int kd, const double* __restrict__ A,
                                                              int lda, int ta, const double* __restrict__ B,
                                                              int ldb, int tb, double* __restrict__ C, int ldc,
                                                              const double* __restrict__ cp_src,
                                                              double* __restrict__ cp_dst) {
  __shared__ SgPanel pa, pb;
  const int i0 = blockIdx.y * SG_T, j0 = blockIdx.x * SG_T;
  sg_panel(pa, A, lda, ta != 0, i0, kd, m);  // pa[k][r] = op(A)(i0 + r, k)
  sg_panel(pb, B, ldb, tb == 0, j0, kd, nc);  // pb[k][c] = op(B)(k, j0 + c)
  __syncthreads();
  const int r = threadIdx.x / SG_T, c = threadIdx.x - r * SG_T;
  const int i = i0 + r, j = j0 + c;
  if (i < m && j < nc) {
    C[(size_t)i * ldc + j] = sg_dot(pa, pb, r, c, kd);
    if (cp_src) cp_dst[(size_t)i * ldc + j] = cp_src[(size_t)i * ldc + j];
  }
}

// F = (H + J) - K and R = F PS - PS^T F (the DIIS commutator residual with
// PS = P_a S), one tile of each per workgroup.
__global__ __launch_bounds__(SG_THREADS) void scf_fock_residual_kernel(int n, const double* __restrict__ H,
                                                                       const double* __restrict__ J,
                                                                       const double* __restrict__ K,
                                                                       const double* __restrict__ PS,
                                                                       double* __restrict__ F,
                                                                       double* __restrict__ R) {
  __shared__ SgPanel fr, psj, psi, fc;
  const int i0 = blockIdx.y * SG_T, j0 = blockIdx.x * SG_T;
  sg_panel_fock(fr, H, J, K, n, false, i0);  // fr[k][r] = F(i0 + r, k)
  sg_panel(psj, PS, n, true, j0, n, n);      // psj[k][c] = PS(k, j0 + c)
  sg_panel(psi, PS, n, true, i0, n, n);      // psi[k][r] = PS(k, i0 + r) = PS^T(i0 + r, k)
  sg_panel_fock(fc, H, J, K, n, true, j0);   // fc[k][c] = F(k, j0 + c)
  __syncthreads();
  const int r = threadIdx.x / SG_T, c = threadIdx.x - r * SG_T;
  const int i = i0 + r, j = j0 + c;
  if (i < n && j < n) {
    const double a = sg_dot(fr, psj, r, c, n), b = sg_dot(psi, fc, r, c, n);
    R[(size_t)i * n + j] = a - b;
    F[(size_t)i * n + j] = fc[i][c];
  }
}

// A = (At + At^T) / 2 with At = XQ^T T (T = F XQ: the Fock matrix in the
// previous orbitals' basis, symmetrised), and V = Q (the eigensolver's warm
// start), one tile of each per workgroup.
__global__ __launch_bounds__(SG_THREADS) void scf_orbital_matrix_kernel(int n, const double* __restrict__ XQ,
                                                                        const double* __restrict__ T,
                                                                        double* __restrict__ A,
                                                                        const double* __restrict__ Q,
                                                                        double* __restrict__ V) {
  __shared__ SgPanel xi, tj, xj, ti;
  const int i0 = blockIdx.y * SG_T, j0 = blockIdx.x * SG_T;
  sg_panel(xi, XQ, n, true, i0, n, n);  // XQ(k, i0 + r)
  sg_panel(tj, T, n, true, j0, n, n);   // T(k, j0 + c)
  sg_panel(xj, XQ, n, true, j0, n, n);  // XQ(k, j0 + c)
  sg_panel(ti, T, n, true, i0, n, n);   // T(k, i0 + r)
  __syncthreads();
  const int r = threadIdx.x / SG_T, c = threadIdx.x - r * SG_T;
  const int i = i0 + r, j = j0 + c;
  if (i < n && j < n) {
    const double aij = sg_dot(xi, tj, r, c, n), aji = sg_dot(ti, xj, r, c, n);
    A[(size_t)i * n + j] = (aij + aji) * 0.5;
    V[(size_t)i * n + j] = Q[(size_t)i * n + j];
  }
}

static dim3 sg_grid(int m, int nc) { return dim3((nc + SG_T - 1) / SG_T, (m + SG_T - 1) / SG_T); }

hipError_t launch_scf_gemm(int m, int nc, int kd, const double* A, int lda, int ta, const double* B, int ldb, int tb,
                           double* C, int ldc, const double* cp_src, double* cp_dst, hipStream_t s) {
  if (m < 1 || nc < 1 || kd < 1 || m > JAC_NMAX || nc > JAC_NMAX || kd > JAC_NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(scf_gemm_kernel, sg_grid(m, nc), dim3(SG_THREADS), 0, s, m, nc, kd, A, lda, ta, B, ldb, tb, C,
                     ldc, cp_src, cp_dst);
  return hipGetLastError();
}

hipError_t launch_scf_fock_residual(int n, const double* H, const double* J, const double* K, const double* PS,
                                    double* F, double* R, hipStream_t s) {
  if (n < 1 || n > JAC_NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(scf_fock_residual_kernel, sg_grid(n, n), dim3(SG_THREADS), 0, s, n, H, J, K, PS, F, R);
  return hipGetLastError();
}

hipError_t launch_scf_orbital_matrix(int n, const double* XQ, const double* T, double* A, const double* Q, double* V,
                                     hipStream_t s) {
  if (n < 1 || n > JAC_NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(scf_orbital_matrix_kernel, sg_grid(n, n), dim3(SG_THREADS), 0, s, n, XQ, T, A, Q, V);
  return hipGetLastError();
}

}  // namespace dft
