/*
 * md_ref.c - C restatement of the reference's McMurchie-Davidson ERI and
 * J/K arithmetic, for the CPU baseline and as a second oracle.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's cpu_baseline
 * leg, never by the dft_amd product.
 *
 * Follows, per primitive Cartesian quartet, exactly the reference device code:
 *   taylor/boys        shelled.py:174-211 (T<1e-14 / 40-term Taylor / erf)
 *   calc_E             shelled.py:228-294
 *   calc_R             shelled.py:296-348 (dense [x][y][z][n] box)
 *   convolution        shelled.py:213-226
 *   eri_kernel body    shelled.py:475-531 (prefactor 2 pi^{5/2}/(p q sqrt(p+q)))
 * the pair records of nuclear_kernel (shelled.py:350-450), normalisation and
 * contraction of calc_eri (shelled.py:663-689) and the eight-orientation,
 * degeneracy-weighted digestion of fill_matrices (shelled.py:791-859), with
 * the add.at scatter replaced by per-thread accumulators summed at the end.
 *
 * Build: make -C oracle   (gcc -O3 -march=native -fopenmp)
 */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAX_E 7
#define MAX_CONV 13
#define MAX_R 2048
#define MAX_BOYS 13

static double taylor(int m, double T) {
  double result = 1.0 / (2.0 * m + 1.0);
  double T_pow = 1.0;
  for (int k = 1; k < 40; ++k) {
    T_pow *= -T / ((double)k);
    result += T_pow / (2.0 * m + 2.0 * k + 1.0);
  }
  return result;
}

void ref_boys(int max_m, double T, double* F) {
  double e = exp(-T);
  if (T < 1e-14) {
    for (int m = 0; m < max_m; ++m) F[m] = 1.0 / (2.0 * m + 1.0);
  } else if (T < 6) {
    double val = taylor(max_m - 1, T);
    F[max_m - 1] = val;
    for (int m = max_m - 2; m >= 0; --m) {
      val = 2.0 * T * val + e;
      val /= 2.0 * m + 1;
      F[m] = val;
    }
  } else {
    double val = sqrt(M_PI) * erf(sqrt(T));
    val /= 2.0 * sqrt(T);
    F[0] = val;
    for (int m = 0; m < max_m - 1; ++m) {
      val = (2.0 * m + 1.0) * val - e;
      val /= 2.0 * T;
      F[m + 1] = val;
    }
  }
}

static void calc_E(double alpha, double beta, double A, double B, int la, int lb, double* E) {
  double prev[MAX_E], next[MAX_E];
  for (int t = 0; t < MAX_E; ++t) prev[t] = next[t] = E[t] = 0.0;
  double p = alpha + beta, q = alpha * beta / p, Q = A - B;
  prev[0] = exp(-q * Q * Q);
  int order = 0;
  for (int step = 0; step < la + lb; ++step) {
    double coef = step < la ? -(beta * Q / p) : (alpha * Q / p);
    int no = order + 1;
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

#define RIDX(x, y, z, n) ((((x) * yl + (y)) * zl + (z)) * nl + (n))
static void calc_R(int xl, int yl, int zl, int nl, double dx, double dy, double dz, double param, double* R) {
  double T = param * (dx * dx + dy * dy + dz * dz);
  double F[MAX_BOYS];
  ref_boys(nl, T, F);
  double scale = 1.0;
  for (int n = 0; n < nl; ++n) {
    R[RIDX(0, 0, 0, n)] = scale * F[n];
    scale *= -2.0 * param;
  }
  for (int x = 0; x < xl; ++x)
    for (int y = 0; y < yl; ++y)
      for (int z = 0; z < zl; ++z) {
        if (x == 0 && y == 0 && z == 0) continue;
        int valid = nl - (x + y + z);
        if (x != 0) {
          for (int n = 0; n < valid; ++n) R[RIDX(x, y, z, n)] = dx * R[RIDX(x - 1, y, z, n + 1)];
          if (x > 1)
            for (int n = 0; n < valid; ++n) R[RIDX(x, y, z, n)] += (x - 1) * R[RIDX(x - 2, y, z, n + 1)];
        } else if (y != 0) {
          for (int n = 0; n < valid; ++n) R[RIDX(x, y, z, n)] = dy * R[RIDX(x, y - 1, z, n + 1)];
          if (y > 1)
            for (int n = 0; n < valid; ++n) R[RIDX(x, y, z, n)] += (y - 1) * R[RIDX(x, y - 2, z, n + 1)];
        } else {
          for (int n = 0; n < valid; ++n) R[RIDX(x, y, z, n)] = dz * R[RIDX(x, y, z - 1, n + 1)];
          if (z > 1)
            for (int n = 0; n < valid; ++n) R[RIDX(x, y, z, n)] += (z - 1) * R[RIDX(x, y, z - 2, n + 1)];
        }
      }
}

static void convolution(const double* Eab, const double* Ecd, int li, int lj, double* C) {
  for (int s = 0; s < li + lj - 1; ++s) C[s] = 0.0;
  for (int t = 0; t < li; ++t)
    for (int tau = 0; tau < lj; ++tau) C[t + tau] += Eab[t] * Ecd[tau] * ((tau & 1) ? -1.0 : 1.0);
}

/* --------------------------------------------------------- pair records */
typedef struct {
  int K;
  double *p, *P, *Ex, *Ey, *Ez;
  int *nx, *ny, *nz;
} Pair;

typedef struct {
  int l, nprim, ncomp, ao;
  double xyz[3];
  const double* exps;
  const double* coef; /* raw contraction coefficients (one row) */
  double* norm;       /* ncomp * nprim, component-major */
  int comp[10][3];
} Shell;

static double dfact(int x) {
  double r = 1.0;
  while (x > 1) { r *= x; x -= 2; }
  return r;
}

static void shell_init(Shell* s) {
  int k = 0;
  for (int n = 0; n <= s->l; ++n)
    for (int m = 0; m <= s->l; ++m)
      for (int l = 0; l <= s->l; ++l)
        if (l + m + n == s->l) { s->comp[k][0] = l; s->comp[k][1] = m; s->comp[k][2] = n; ++k; }
  s->ncomp = k;
  s->norm = (double*)malloc(sizeof(double) * k * s->nprim);
  for (int c = 0; c < k; ++c)
    for (int i = 0; i < s->nprim; ++i) {
      double a = s->exps[i];
      double t = pow(4 * a, s->l) / (dfact(2 * s->comp[c][0] - 1) * dfact(2 * s->comp[c][1] - 1) *
                                     dfact(2 * s->comp[c][2] - 1));
      s->norm[c * s->nprim + i] = sqrt(t) * pow(2 * a / M_PI, 0.75);
    }
}

static void pair_init(const Shell* A, const Shell* B, Pair* P) {
  int na = A->ncomp * A->nprim, nb = B->ncomp * B->nprim;
  P->K = na * nb;
  P->p = malloc(sizeof(double) * P->K);
  P->P = malloc(sizeof(double) * 3 * P->K);
  P->Ex = malloc(sizeof(double) * MAX_E * P->K);
  P->Ey = malloc(sizeof(double) * MAX_E * P->K);
  P->Ez = malloc(sizeof(double) * MAX_E * P->K);
  P->nx = malloc(sizeof(int) * P->K);
  P->ny = malloc(sizeof(int) * P->K);
  P->nz = malloc(sizeof(int) * P->K);
  for (int q = 0; q < P->K; ++q) {
    int a = q / nb, b = q % nb;
    int ca = a / A->nprim, ia = a % A->nprim, cb = b / B->nprim, ib = b % B->nprim;
    double e1 = A->exps[ia], e2 = B->exps[ib];
    calc_E(e1, e2, A->xyz[0], B->xyz[0], A->comp[ca][0], B->comp[cb][0], P->Ex + q * MAX_E);
    calc_E(e1, e2, A->xyz[1], B->xyz[1], A->comp[ca][1], B->comp[cb][1], P->Ey + q * MAX_E);
    calc_E(e1, e2, A->xyz[2], B->xyz[2], A->comp[ca][2], B->comp[cb][2], P->Ez + q * MAX_E);
    double p = e1 + e2;
    P->p[q] = p;
    for (int d = 0; d < 3; ++d) P->P[3 * q + d] = (e1 * A->xyz[d] + e2 * B->xyz[d]) / p;
    P->nx[q] = A->comp[ca][0] + B->comp[cb][0] + 1;
    P->ny[q] = A->comp[ca][1] + B->comp[cb][1] + 1;
    P->nz[q] = A->comp[ca][2] + B->comp[cb][2] + 1;
  }
}

static void pair_free(Pair* P) {
  free(P->p); free(P->P); free(P->Ex); free(P->Ey); free(P->Ez); free(P->nx); free(P->ny); free(P->nz);
}

/* one eri_kernel thread: primitive Cartesian (i | j) */
static double eri_prim(const Pair* ab, int i, const Pair* cd, int j, double* R) {
  double pi_ = ab->p[i], pj = cd->p[j];
  int xl = ab->nx[i] + cd->nx[j] - 1, yl = ab->ny[i] + cd->ny[j] - 1, zl = ab->nz[i] + cd->nz[j] - 1;
  int nl = xl + yl + zl - 2;
  double rho = pi_ * pj / (pi_ + pj);
  calc_R(xl, yl, zl, nl, ab->P[3 * i] - cd->P[3 * j], ab->P[3 * i + 1] - cd->P[3 * j + 1],
         ab->P[3 * i + 2] - cd->P[3 * j + 2], rho, R);
  double Cx[MAX_CONV], Cy[MAX_CONV], Cz[MAX_CONV];
  convolution(ab->Ex + i * MAX_E, cd->Ex + j * MAX_E, ab->nx[i], cd->nx[j], Cx);
  convolution(ab->Ey + i * MAX_E, cd->Ey + j * MAX_E, ab->ny[i], cd->ny[j], Cy);
  convolution(ab->Ez + i * MAX_E, cd->Ez + j * MAX_E, ab->nz[i], cd->nz[j], Cz);
  double s = 0.0;
  for (int x = 0; x < xl; ++x)
    for (int y = 0; y < yl; ++y) {
      double cxy = Cx[x] * Cy[y];
      for (int z = 0; z < zl; ++z) s += cxy * Cz[z] * R[RIDX(x, y, z, 0)];
    }
  return s * (2.0 * M_PI * M_PI * sqrt(M_PI)) / (pi_ * pj * sqrt(pi_ + pj));
}

/* contracted, normalised block [ncomp_a][ncomp_b][ncomp_c][ncomp_d] */
static void shell_quartet(const Shell* S[4], const Pair* ab, const Pair* cd, double* prim, double* R, double* out) {
  for (int i = 0; i < ab->K; ++i)
    for (int j = 0; j < cd->K; ++j) prim[(long)i * cd->K + j] = eri_prim(ab, i, cd, j, R);
  int nc[4], np[4];
  for (int k = 0; k < 4; ++k) { nc[k] = S[k]->ncomp; np[k] = S[k]->nprim; }
  int nb_ = nc[1] * np[1], nc_ = nc[2] * np[2], nd_ = nc[3] * np[3];
  for (int a = 0; a < nc[0]; ++a)
    for (int b = 0; b < nc[1]; ++b)
      for (int c = 0; c < nc[2]; ++c)
        for (int d = 0; d < nc[3]; ++d) {
          double s = 0.0;
          for (int p = 0; p < np[0]; ++p) {
            int ia = a * np[0] + p;
            double wa = S[0]->norm[ia] * S[0]->coef[p];
            for (int q = 0; q < np[1]; ++q) {
              int ib = b * np[1] + q;
              double wb = wa * S[1]->norm[ib] * S[1]->coef[q];
              long row = (long)ia * nb_ + ib;
              for (int r = 0; r < np[2]; ++r) {
                int ic = c * np[2] + r;
                double wc = wb * S[2]->norm[ic] * S[2]->coef[r];
                for (int t = 0; t < np[3]; ++t) {
                  int id = d * np[3] + t;
                  s += prim[row * (nc_ * nd_) + (long)ic * nd_ + id] * wc * S[3]->norm[id] * S[3]->coef[t];
                }
              }
            }
          }
          out[((a * nc[1] + b) * nc[2] + c) * nc[3] + d] = s;
        }
}

typedef struct {
  int nsh, nbf;
  long maxK;
  Shell* sh;
  Pair* pairs; /* nsh*nsh, lower triangle used */
} Ctx;

static Ctx* ctx_new(int nsh, const int* l, const int* nprim, const int* ao, const double* xyz, const double* exps,
                    const double* coefs, int nbf) {
  Ctx* c = calloc(1, sizeof(Ctx));
  c->nsh = nsh;
  c->nbf = nbf;
  c->sh = calloc(nsh, sizeof(Shell));
  int off = 0;
  for (int i = 0; i < nsh; ++i) {
    Shell* s = &c->sh[i];
    s->l = l[i];
    s->nprim = nprim[i];
    s->ao = ao[i];
    for (int d = 0; d < 3; ++d) s->xyz[d] = xyz[3 * i + d];
    s->exps = exps + off;
    s->coef = coefs + off;
    off += nprim[i];
    shell_init(s);
  }
  c->pairs = calloc((size_t)nsh * nsh, sizeof(Pair));
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j <= i; ++j) {
      pair_init(&c->sh[i], &c->sh[j], &c->pairs[i * nsh + j]);
      if (c->pairs[i * nsh + j].K > c->maxK) c->maxK = c->pairs[i * nsh + j].K;
    }
  return c;
}

static void ctx_free(Ctx* c) {
  for (int i = 0; i < c->nsh; ++i) {
    free(c->sh[i].norm);
    for (int j = 0; j <= i; ++j) pair_free(&c->pairs[i * c->nsh + j]);
  }
  free(c->pairs);
  free(c->sh);
  free(c);
}

/* blocks for quartets (a>=b, c>=d as in calc_eri); out concatenated */
int ref_eri_blocks(int nsh, const int* l, const int* nprim, const int* ao, const double* xyz, const double* exps,
                   const double* coefs, int nbf, const int* quartets, long long nq, double* out) {
  Ctx* c = ctx_new(nsh, l, nprim, ao, xyz, exps, coefs, nbf);
  long long* offs = malloc(sizeof(long long) * (nq + 1));
  offs[0] = 0;
  for (long long q = 0; q < nq; ++q) {
    long long sz = 1;
    for (int k = 0; k < 4; ++k) sz *= c->sh[quartets[4 * q + k]].ncomp;
    offs[q + 1] = offs[q] + sz;
  }
#pragma omp parallel
  {
    double* R = malloc(sizeof(double) * MAX_R);
    double* prim = malloc(sizeof(double) * c->maxK * c->maxK);
#pragma omp for schedule(dynamic, 16)
    for (long long q = 0; q < nq; ++q) {
      const int* s = quartets + 4 * q;
      const Shell* S[4] = {&c->sh[s[0]], &c->sh[s[1]], &c->sh[s[2]], &c->sh[s[3]]};
      shell_quartet(S, &c->pairs[s[0] * nsh + s[1]], &c->pairs[s[2] * nsh + s[3]], prim, R, out + offs[q]);
    }
    free(prim);
    free(R);
  }
  free(offs);
  ctx_free(c);
  return 0;
}

/*
 * Direct J/K over a list of unique quartets (a>=b, c>=d, ab>=cd): the eight
 * orientations of fill_matrices with weight 1/2^([a=b]+[c=d]+[ab=cd]),
 * accumulated per thread and summed. J, Ka, Kb are overwritten. Pb/Kb may be
 * NULL. Returns 0.
 */
int ref_build_jk(int nsh, const int* l, const int* nprim, const int* ao, const double* xyz, const double* exps,
                 const double* coefs, int nbf, const int* quartets, long long nq, const double* P,
                 const double* Pa, const double* Pb, double* J, double* Ka, double* Kb, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  Ctx* c = ctx_new(nsh, l, nprim, ao, xyz, exps, coefs, nbf);
  const long nn = (long)nbf * nbf;
  int nt = omp_get_max_threads();
  double* acc = calloc((size_t)nt * 3 * nn, sizeof(double));
#pragma omp parallel
  {
    int tid = omp_get_thread_num();
    double* Jt = acc + (size_t)tid * 3 * nn;
    double* Kat = Jt + nn;
    double* Kbt = Kat + nn;
    double* R = malloc(sizeof(double) * MAX_R);
    double* prim = malloc(sizeof(double) * c->maxK * c->maxK);
    double* blk = malloc(sizeof(double) * 10000);
#pragma omp for schedule(dynamic, 64)
    for (long long q = 0; q < nq; ++q) {
      const int* s = quartets + 4 * q;
      const Shell* S[4] = {&c->sh[s[0]], &c->sh[s[1]], &c->sh[s[2]], &c->sh[s[3]]};
      shell_quartet(S, &c->pairs[s[0] * nsh + s[1]], &c->pairs[s[2] * nsh + s[3]], prim, R, blk);
      int n[4], o[4];
      for (int k = 0; k < 4; ++k) { n[k] = S[k]->ncomp; o[k] = S[k]->ao; }
      double w = 1.0 / (double)(1 << ((s[0] == s[1]) + (s[2] == s[3]) + (s[0] == s[2] && s[1] == s[3])));
      static const int orient[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                       {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
      for (int r = 0; r < 8; ++r) {
        const int* sw = orient[r];
        int x[4];
        for (x[0] = 0; x[0] < n[0]; ++x[0])
          for (x[1] = 0; x[1] < n[1]; ++x[1])
            for (x[2] = 0; x[2] < n[2]; ++x[2])
              for (x[3] = 0; x[3] < n[3]; ++x[3]) {
                double v = w * blk[((x[0] * n[1] + x[1]) * n[2] + x[2]) * n[3] + x[3]];
                /* orientation sw: (i j | k l) = positions sw[0..3] of the block */
                int i = o[sw[0]] + x[sw[0]], j = o[sw[1]] + x[sw[1]];
                int k = o[sw[2]] + x[sw[2]], m = o[sw[3]] + x[sw[3]];
                Jt[(long)i * nbf + j] += v * P[(long)k * nbf + m];
                Kat[(long)i * nbf + k] += v * Pa[(long)j * nbf + m];
                if (Pb) Kbt[(long)i * nbf + k] += v * Pb[(long)j * nbf + m];
              }
      }
    }
    free(blk);
    free(prim);
    free(R);
  }
  for (long e = 0; e < nn; ++e) {
    double sj = 0, sa = 0, sb = 0;
    for (int t = 0; t < nt; ++t) {
      sj += acc[(size_t)t * 3 * nn + e];
      sa += acc[(size_t)t * 3 * nn + nn + e];
      sb += acc[(size_t)t * 3 * nn + 2 * nn + e];
    }
    J[e] = sj;
    Ka[e] = sa;
    if (Kb) Kb[e] = sb;
  }
  free(acc);
  ctx_free(c);
  return 0;
}

int ref_max_threads(void) { return omp_get_max_threads(); }
