/*
 * oracle.c -- CPU restatement of the reference local contraction path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that every product and sum is rounded separately, like the reference's
 * x86-64 build of std::complex<double> arithmetic (no FMA without -march).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define MAXR 64

uint64_t or_size(int rank, const uint64_t *dims) {
  uint64_t n = 1;
  for (int i = 0; i < rank; ++i) n *= dims[i];
  return n;
}

static void strides_of(int rank, const uint64_t *dims, uint64_t *st) {
  uint64_t s = 1;
  for (int i = rank - 1; i >= 0; --i) { st[i] = s; s *= dims[i]; }
}

/* ---------------------------------------------------------------------------
 * RNG.  SplitMix64: state_k = seed + (k+1)*0x9E3779B97F4A7C15, output = mix(state_k).
 * --------------------------------------------------------------------------- */
uint64_t or_splitmix64(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double bits_to_double(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t double_to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

double or_log(double u) {
  /* u = m * 2^e with m in [sqrt(1/2), sqrt(2)); log u = e*ln2 + 2*atanh((m-1)/(m+1)). */
  uint64_t b = double_to_bits(u);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  double p = 0x1.47ae147ae147bp-5;          /* 1/25 */
  p = p * z + 0x1.642c8590b2164p-5;         /* 1/23 */
  p = p * z + 0x1.8618618618618p-5;         /* 1/21 */
  p = p * z + 0x1.af286bca1af28p-5;         /* 1/19 */
  p = p * z + 0x1.e1e1e1e1e1e1ep-5;         /* 1/17 */
  p = p * z + 0x1.1111111111111p-4;         /* 1/15 */
  p = p * z + 0x1.3b13b13b13b14p-4;         /* 1/13 */
  p = p * z + 0x1.745d1745d1746p-4;         /* 1/11 */
  p = p * z + 0x1.c71c71c71c71cp-4;         /* 1/9 */
  p = p * z + 0x1.2492492492492p-3;         /* 1/7 */
  p = p * z + 0x1.999999999999ap-3;         /* 1/5 */
  p = p * z + 0x1.5555555555555p-2;         /* 1/3 */
  double at = s + s * (z * p);              /* atanh(s) */
  double de = (double)e;
  return de * 0x1.62e42fefa39efp-1 + (de * 0x1.abc9e3b39803fp-56 + 2.0 * at);
}

static void sincos_poly(double a, double *s, double *c) {
  double z = a * a;
  double ps = -0x1.2f49b46814157p-57;
  ps = ps * z + 0x1.952c77030ad4ap-49;
  ps = ps * z + -0x1.ae7f3e733b81fp-41;
  ps = ps * z + 0x1.6124613a86d09p-33;
  ps = ps * z + -0x1.ae64567f544e4p-26;
  ps = ps * z + 0x1.71de3a556c734p-19;
  ps = ps * z + -0x1.a01a01a01a01ap-13;
  ps = ps * z + 0x1.1111111111111p-7;
  ps = ps * z + -0x1.5555555555555p-3;
  *s = a + a * (z * ps);
  double pc = -0x1.6827863b97d97p-53;
  pc = pc * z + 0x1.ae7f3e733b81fp-45;
  pc = pc * z + -0x1.93974a8c07c9dp-37;
  pc = pc * z + 0x1.1eed8eff8d898p-29;
  pc = pc * z + -0x1.27e4fb7789f5cp-22;
  pc = pc * z + 0x1.a01a01a01a01ap-16;
  pc = pc * z + -0x1.6c16c16c16c17p-10;
  pc = pc * z + 0x1.5555555555555p-5;
  pc = pc * z + -0x1.0000000000000p-1;
  *c = 1.0 + z * pc;
}

void or_sincos2pi(double v, double *s, double *c) {
  double t = v * 8.0;                 /* exact */
  int o = (int)t;                     /* octant 0..7 */
  double f = t - (double)o;           /* exact, in [0, 1) */
  if (o & 1) f = 1.0 - f;             /* exact */
  double a = f * 0x1.921fb54442d18p-1; /* f * pi/4 */
  double ss, cc;
  sincos_poly(a, &ss, &cc);
  switch (o & 7) {
    case 0: *s = ss;  *c = cc;  break;
    case 1: *s = cc;  *c = ss;  break;
    case 2: *s = cc;  *c = -ss; break;
    case 3: *s = ss;  *c = -cc; break;
    case 4: *s = -ss; *c = -cc; break;
    case 5: *s = -cc; *c = -ss; break;
    case 6: *s = -cc; *c = ss;  break;
    default: *s = -ss; *c = cc; break;
  }
}

void or_fill_c128(double *out, uint64_t first, uint64_t n, uint64_t seed, int dist) {
  const double two_m53 = 0x1.0p-53;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t g = first + i;
    uint64_t r0 = or_splitmix64(seed, 2 * g);
    uint64_t r1 = or_splitmix64(seed, 2 * g + 1);
    if (dist == 1) {
      out[2 * i] = (double)(r0 >> 11) * two_m53 * 2.0 - 1.0;
      out[2 * i + 1] = (double)(r1 >> 11) * two_m53 * 2.0 - 1.0;
    } else {
      double u1 = (double)((r0 >> 11) + 1) * two_m53; /* (0, 1] */
      double u2 = (double)(r1 >> 11) * two_m53;       /* [0, 1) */
      double rad = __builtin_sqrt(-2.0 * or_log(u1));
      double sn, cs;
      or_sincos2pi(u2, &sn, &cs);
      out[2 * i] = rad * cs;
      out[2 * i + 1] = rad * sn;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Permute and swap.
 * --------------------------------------------------------------------------- */
int or_permute(const double *src, double *dst, int rank, const uint64_t *dims, const uint16_t *ptup) {
  if (rank < 0 || rank > MAXR) return OR_EINVAL;
  int seen[MAXR] = {0};
  for (int i = 0; i < rank; ++i) {
    if (ptup[i] >= rank || seen[ptup[i]]) return OR_EINVAL;
    seen[ptup[i]] = 1;
  }
  uint64_t nd[MAXR], nst[MAXR], jump[MAXR];
  for (int i = 0; i < rank; ++i) nd[ptup[i]] = dims[i];
  strides_of(rank, nd, nst);
  for (int i = 0; i < rank; ++i) jump[i] = nst[ptup[i]];
  uint64_t n = or_size(rank, dims);
  uint64_t idx[MAXR] = {0};
  uint64_t off = 0;
  for (uint64_t s = 0; s < n; ++s) {
    dst[2 * off] = src[2 * s];
    dst[2 * off + 1] = src[2 * s + 1];
    for (int k = rank - 1; k >= 0; --k) {
      if (++idx[k] < dims[k]) { off += jump[k]; break; }
      off -= (dims[k] - 1) * jump[k];
      idx[k] = 0;
    }
  }
  return OR_OK;
}

int or_swap(double *data, int rank, const uint64_t *dims, int i, int j) {
  if (i < 0 || j < 0 || i >= rank || j >= rank) return OR_EINVAL;
  if (i > j) { int t = i; i = j; j = t; }
  if (dims[i] != dims[j]) return OR_EINVAL;
  if (i == j) return OR_OK;
  uint64_t st[MAXR];
  strides_of(rank, dims, st);
  uint64_t n = or_size(rank, dims);
  /* Walk all elements; exchange each pair (a at i, b at j) with a < b once. */
  uint64_t idx[MAXR] = {0};
  for (uint64_t s = 0; s < n; ++s) {
    if (idx[i] < idx[j]) {
      uint64_t o = s - idx[i] * st[i] - idx[j] * st[j] + idx[j] * st[i] + idx[i] * st[j];
      double r = data[2 * s], m = data[2 * s + 1];
      data[2 * s] = data[2 * o]; data[2 * s + 1] = data[2 * o + 1];
      data[2 * o] = r; data[2 * o + 1] = m;
    }
    for (int k = rank - 1; k >= 0; --k) {
      if (++idx[k] < dims[k]) break;
      idx[k] = 0;
    }
  }
  return OR_OK;
}

/* ---------------------------------------------------------------------------
 * Output placement.
 * --------------------------------------------------------------------------- */
static int wired_as(const uint16_t *wires, int nw, int side, int axis) {
  for (int w = 0; w < nw; ++w) if (wires[2 * w + side] == axis) return w;
  return -1;
}

int or_default_repls(int r1, int ndis1, int r2, int ndis2, int nw, const uint16_t *wires,
                     uint16_t *repl1, uint16_t *repl2) {
  int ndw = 0;
  for (int w = 0; w < nw; ++w) if (wires[2 * w] < ndis1) ++ndw;
  /* t1: open axes keep relative order; local ones shift past t2's open distributed axes. */
  int skipped = 0;
  for (int i = 0; i < r1; ++i) {
    if (wired_as(wires, nw, 0, i) >= 0) { repl1[i] = OR_X; ++skipped; continue; }
    int p = i - skipped;
    if (i >= ndis1) p += ndis2 - ndw;
    repl1[i] = (uint16_t)p;
  }
  skipped = 0;
  for (int i = 0; i < r2; ++i) {
    if (wired_as(wires, nw, 1, i) >= 0) { repl2[i] = OR_X; ++skipped; continue; }
    int p = (i < ndis2) ? (ndis1 - ndw + i - skipped) : (r1 - nw + i - skipped);
    repl2[i] = (uint16_t)p;
  }
  return OR_OK;
}

int or_symm_repls(int r1, int nd2, int nl2, int n_dis_in, int nw, const uint16_t *wires,
                  uint16_t *repl1, uint16_t *repl2) {
  /* Local input axes: the first nl2/2 - n_dis_in local axes (symm.hpp:26-28).  The
   * reference leaves a negative count undefined; it is read as zero here. */
  int n_loc_in = nl2 / 2 - n_dis_in;
  if (n_loc_in < 0) n_loc_in = 0;
  int inputs = 0;
  for (int w = 0; w < nw; ++w) {
    int b = wires[2 * w + 1];
    if (b < n_dis_in || (b >= nd2 && b < nd2 + n_loc_in)) ++inputs;
  }
  int imbal = n_dis_in - (nd2 - n_dis_in);
  if (imbal != 0 || inputs != n_dis_in + n_loc_in) return 0;
  for (int i = 0; i < r1; ++i) repl1[i] = (wired_as(wires, nw, 0, i) >= 0) ? OR_X : (uint16_t)i;
  /* Wires ordered by their t2 axis give "from" positions; the k-th open axis of
   * t2 (ascending) takes the k-th of them. */
  int order[MAXR];
  for (int w = 0; w < nw; ++w) order[w] = w;
  for (int a = 1; a < nw; ++a)
    for (int b = a; b > 0 && wires[2 * order[b] + 1] < wires[2 * order[b - 1] + 1]; --b) {
      int t = order[b]; order[b] = order[b - 1]; order[b - 1] = t;
    }
  int r2 = nd2 + nl2, k = 0;
  for (int i = 0; i < r2; ++i) {
    if (wired_as(wires, nw, 1, i) >= 0) { repl2[i] = OR_X; continue; }
    repl2[i] = (k < nw) ? wires[2 * order[k++]] : OR_X;
  }
  return 1;
}

/* ---------------------------------------------------------------------------
 * Contraction.
 * --------------------------------------------------------------------------- */
int or_contract(const double *A, int rA, const uint64_t *dA,
                const double *B, int rB, const uint64_t *dB,
                int nw, const uint16_t *wires,
                const uint16_t *repl1, const uint16_t *repl2, double *C) {
  if (rA > MAXR || rB > MAXR || nw < 0 || nw > rA || nw > rB) return OR_EINVAL;
  int usedA[MAXR] = {0}, usedB[MAXR] = {0};
  for (int w = 0; w < nw; ++w) {
    int a = wires[2 * w], b = wires[2 * w + 1];
    if (a >= rA || b >= rB || usedA[a] || usedB[b]) return OR_EINVAL;
    if (dA[a] != dB[b]) return OR_EINVAL;
    usedA[a] = usedB[b] = 1;
  }
  /* Closed axes in the reference order: wires sorted by their t2 axis. */
  int order[MAXR];
  for (int w = 0; w < nw; ++w) order[w] = w;
  for (int a = 1; a < nw; ++a)
    for (int b = a; b > 0 && wires[2 * order[b] + 1] < wires[2 * order[b - 1] + 1]; --b) {
      int t = order[b]; order[b] = order[b - 1]; order[b - 1] = t;
    }
  uint64_t stA[MAXR], stB[MAXR];
  strides_of(rA, dA, stA);
  strides_of(rB, dB, stB);
  uint64_t kd[MAXR], kA[MAXR], kB[MAXR];
  uint64_t K = 1;
  for (int w = 0; w < nw; ++w) {
    int a = wires[2 * order[w]], b = wires[2 * order[w] + 1];
    kd[w] = dA[a]; kA[w] = stA[a]; kB[w] = stB[b];
    K *= kd[w];
  }
  /* Natural output axes: open t1 axes then open t2 axes. */
  int rC = rA + rB - 2 * nw;
  if (rC > MAXR) return OR_EINVAL;
  uint64_t od[MAXR], oSrc[MAXR], dC[MAXR];
  int pos[MAXR], isB[MAXR];
  int seen[MAXR] = {0};
  int m = 0;
  for (int i = 0; i < rA; ++i) if (!usedA[i]) {
    if (repl1[i] >= rC || seen[repl1[i]]) return OR_EINVAL;
    seen[repl1[i]] = 1; od[m] = dA[i]; oSrc[m] = stA[i]; pos[m] = repl1[i]; isB[m] = 0; ++m;
  }
  for (int i = 0; i < rB; ++i) if (!usedB[i]) {
    if (repl2[i] >= rC || seen[repl2[i]]) return OR_EINVAL;
    seen[repl2[i]] = 1; od[m] = dB[i]; oSrc[m] = stB[i]; pos[m] = repl2[i]; isB[m] = 1; ++m;
  }
  for (int q = 0; q < rC; ++q) dC[pos[q]] = od[q];
  uint64_t stC[MAXR];
  strides_of(rC, dC, stC);
  uint64_t N = or_size(rC, od);
  uint64_t oi[MAXR] = {0};
  uint64_t offA = 0, offB = 0, offC = 0;
  for (uint64_t e = 0; e < N; ++e) {
    double accr = 0.0, acci = 0.0;
    uint64_t ki[MAXR] = {0};
    uint64_t ka = offA, kb = offB;
    for (uint64_t t = 0; t < K; ++t) {
      double ar = A[2 * ka], ai = A[2 * ka + 1];
      double br = B[2 * kb], bi = B[2 * kb + 1];
      double pr = ar * br - ai * bi;
      double pi = ar * bi + ai * br;
      accr = accr + pr;
      acci = acci + pi;
      for (int w = nw - 1; w >= 0; --w) {
        if (++ki[w] < kd[w]) { ka += kA[w]; kb += kB[w]; break; }
        ka -= (kd[w] - 1) * kA[w]; kb -= (kd[w] - 1) * kB[w];
        ki[w] = 0;
      }
    }
    C[2 * offC] = accr;
    C[2 * offC + 1] = acci;
    for (int q = rC - 1; q >= 0; --q) {
      if (++oi[q] < od[q]) {
        if (isB[q]) offB += oSrc[q]; else offA += oSrc[q];
        offC += stC[pos[q]];
        break;
      }
      if (isB[q]) offB -= (od[q] - 1) * oSrc[q]; else offA -= (od[q] - 1) * oSrc[q];
      offC -= (od[q] - 1) * stC[pos[q]];
      oi[q] = 0;
    }
  }
  return OR_OK;
}

int or_gate_apply(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                  const double *G) {
  if (k <= 0 || k > rank || rank > MAXR) return OR_EINVAL;
  uint64_t gd[MAXR];
  uint16_t wires[2 * MAXR];
  for (int m = 0; m < k; ++m) {
    if (axes[m] >= rank) return OR_EINVAL;
    gd[m] = gd[m + k] = dims[axes[m]];
    wires[2 * m] = axes[m];
    wires[2 * m + 1] = (uint16_t)m;
  }
  uint16_t repl1[MAXR], repl2[MAXR];
  if (!or_symm_repls(rank, 0, 2 * k, 0, k, wires, repl1, repl2)) return OR_EINVAL;
  uint64_t n = or_size(rank, dims);
  double *tmp = (double *)malloc(sizeof(double) * 2 * (n ? n : 1));
  if (!tmp) return OR_ENOMEM;
  int rc = or_contract(state, rank, dims, G, 2 * k, gd, k, wires, repl1, repl2, tmp);
  if (rc == OR_OK) memcpy(state, tmp, sizeof(double) * 2 * n);
  free(tmp);
  return rc;
}

int or_diag_apply(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                  const double *d) {
  if (k <= 0 || k > rank || rank > MAXR) return OR_EINVAL;
  uint64_t st[MAXR], dst[MAXR];
  strides_of(rank, dims, st);
  uint64_t s = 1;
  for (int m = k - 1; m >= 0; --m) {
    if (axes[m] >= rank) return OR_EINVAL;
    dst[m] = s; s *= dims[axes[m]];
  }
  uint64_t n = or_size(rank, dims);
  for (uint64_t e = 0; e < n; ++e) {
    uint64_t di = 0;
    for (int m = 0; m < k; ++m) di += ((e / st[axes[m]]) % dims[axes[m]]) * dst[m];
    double ar = state[2 * e], ai = state[2 * e + 1];
    double br = d[2 * di], bi = d[2 * di + 1];
    state[2 * e] = ar * br - ai * bi;
    state[2 * e + 1] = ar * bi + ai * br;
  }
  return OR_OK;
}

int or_qft(double *state, int n, const double *phases) {
  uint64_t dims[MAXR];
  if (n <= 0 || n > MAXR) return OR_EINVAL;
  for (int i = 0; i < n; ++i) dims[i] = 2;
  const double h = 0x1.6a09e667f3bcdp-1; /* 2^-1/2 as pow(2, -.5) gives */
  const double H[8] = {h, 0, h, 0, h, 0, -h, 0};
  for (int i = 0; i < n; ++i) {
    uint16_t ax = (uint16_t)i;
    int rc = or_gate_apply(state, n, dims, 1, &ax, H);
    if (rc) return rc;
    for (int j = i + 1; j < n; ++j) {
      double d[8] = {1, 0, 1, 0, 1, 0, phases[2 * (j - i)], phases[2 * (j - i) + 1]};
      uint16_t axes[2] = {(uint16_t)i, (uint16_t)j};
      rc = or_diag_apply(state, n, dims, 2, axes, d);
      if (rc) return rc;
    }
  }
  return OR_OK;
}

double or_norm2(const double *x, uint64_t n) {
  double s = 0.0;
  for (uint64_t i = 0; i < 2 * n; ++i) s += x[i] * x[i];
  return s;
}

uint64_t or_multiset_hash(const double *x, uint64_t n) {
  uint64_t h = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t a = double_to_bits(x[2 * i]), b = double_to_bits(x[2 * i + 1]);
    h += or_splitmix64(a ^ (b * 0xD6E8FEB86659FD93ull), 0);
  }
  return h;
}

/* ---------------------------------------------------------------------------
 * Multi-threaded restatements for the CPU timing baseline (not checkers).
 * Each thread takes a contiguous range of elements (permute, diagonal) or of
 * gate groups and does exactly the per-element arithmetic of the functions
 * above, so the results are equal to theirs bit for bit.
 * --------------------------------------------------------------------------- */
#include <pthread.h>

#define MAXTHREADS 256
typedef void (*range_fn)(void *ctx, uint64_t begin, uint64_t end);
struct range_job {
  range_fn fn;
  void *ctx;
  uint64_t b, e;
};

static void *range_thread(void *p) {
  struct range_job *j = (struct range_job *)p;
  j->fn(j->ctx, j->b, j->e);
  return NULL;
}

/* Splits [0, n) into nthreads contiguous ranges; a thread that cannot be
 * started runs its range on the calling thread. */
static void parallel_range(uint64_t n, int nthreads, range_fn fn, void *ctx) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > MAXTHREADS) nthreads = MAXTHREADS;
  pthread_t th[MAXTHREADS];
  int started[MAXTHREADS] = {0};
  struct range_job jobs[MAXTHREADS];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
    jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
  }
  for (int t = 1; t < nthreads; ++t) started[t] = pthread_create(&th[t], NULL, range_thread, &jobs[t]) == 0;
  range_thread(&jobs[0]);
  for (int t = 1; t < nthreads; ++t) {
    if (started[t])
      pthread_join(th[t], NULL);
    else
      range_thread(&jobs[t]);
  }
}

/* Threads own contiguous result ranges (their writes never share a cache
 * line); each result element is read from its source position. */
struct permute_ctx {
  const double *src;
  double *dst;
  int rank;
  uint64_t nd[MAXR];    /* result dims */
  uint64_t jump[MAXR];  /* source stride of result axis p */
};

static void permute_range(void *p, uint64_t b, uint64_t e) {
  const struct permute_ctx *c = (const struct permute_ctx *)p;
  uint64_t idx[MAXR] = {0}, off = 0, rest = b;
  for (int k = c->rank - 1; k >= 0; --k) {
    idx[k] = rest % c->nd[k];
    rest /= c->nd[k];
    off += idx[k] * c->jump[k];
  }
  for (uint64_t t = b; t < e; ++t) {
    c->dst[2 * t] = c->src[2 * off];
    c->dst[2 * t + 1] = c->src[2 * off + 1];
    for (int k = c->rank - 1; k >= 0; --k) {
      if (++idx[k] < c->nd[k]) { off += c->jump[k]; break; }
      off -= (c->nd[k] - 1) * c->jump[k];
      idx[k] = 0;
    }
  }
}

int or_permute_mt(const double *src, double *dst, int rank, const uint64_t *dims, const uint16_t *ptup, int nthreads) {
  if (rank < 0 || rank > MAXR) return OR_EINVAL;
  int seen[MAXR] = {0};
  for (int i = 0; i < rank; ++i) {
    if (ptup[i] >= rank || seen[ptup[i]]) return OR_EINVAL;
    seen[ptup[i]] = 1;
  }
  struct permute_ctx c;
  uint64_t st[MAXR];
  strides_of(rank, dims, st);
  for (int i = 0; i < rank; ++i) {
    c.nd[ptup[i]] = dims[i];
    c.jump[ptup[i]] = st[i];
  }
  c.src = src;
  c.dst = dst;
  c.rank = rank;
  parallel_range(or_size(rank, dims), nthreads, permute_range, &c);
  return OR_OK;
}

struct gate_ctx {
  double *state;
  const double *G;
  int k, nrest;
  uint64_t K;
  uint64_t rest_dims[MAXR], rest_st[MAXR];  /* the non-gate axes, in order */
  uint64_t gdims[MAXR], gst[MAXR];          /* the gate axes, in gate order */
};

static void gate_range(void *p, uint64_t b, uint64_t e) {
  const struct gate_ctx *c = (const struct gate_ctx *)p;
  const uint64_t K = c->K;
  uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * K);
  double *in = (double *)malloc(sizeof(double) * 2 * K);
  if (!off || !in) {
    free(off);
    free(in);
    return;
  }
  /* Offsets of the gate inputs i (row-major over the gate axes, last fastest). */
  for (uint64_t i = 0; i < K; ++i) {
    uint64_t r = i, o = 0;
    for (int m = c->k - 1; m >= 0; --m) {
      o += (r % c->gdims[m]) * c->gst[m];
      r /= c->gdims[m];
    }
    off[i] = o;
  }
  for (uint64_t g = b; g < e; ++g) {
    uint64_t r = g, base = 0;
    for (int q = c->nrest - 1; q >= 0; --q) {
      base += (r % c->rest_dims[q]) * c->rest_st[q];
      r /= c->rest_dims[q];
    }
    for (uint64_t i = 0; i < K; ++i) {
      in[2 * i] = c->state[2 * (base + off[i])];
      in[2 * i + 1] = c->state[2 * (base + off[i]) + 1];
    }
    for (uint64_t o = 0; o < K; ++o) {
      double accr = 0.0, acci = 0.0;
      for (uint64_t i = 0; i < K; ++i) {
        double ar = in[2 * i], ai = in[2 * i + 1];
        double br = c->G[2 * (i * K + o)], bi = c->G[2 * (i * K + o) + 1];
        double pr = ar * br - ai * bi;
        double pi = ar * bi + ai * br;
        accr = accr + pr;
        acci = acci + pi;
      }
      c->state[2 * (base + off[o])] = accr;
      c->state[2 * (base + off[o]) + 1] = acci;
    }
  }
  free(off);
  free(in);
}

int or_gate_apply_mt(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes, const double *G,
                     int nthreads) {
  if (k <= 0 || k > rank || rank > MAXR) return OR_EINVAL;
  struct gate_ctx c;
  uint64_t st[MAXR];
  int is_gate[MAXR] = {0};
  strides_of(rank, dims, st);
  c.K = 1;
  for (int m = 0; m < k; ++m) {
    if (axes[m] >= rank || is_gate[axes[m]]) return OR_EINVAL;
    is_gate[axes[m]] = 1;
    c.gdims[m] = dims[axes[m]];
    c.gst[m] = st[axes[m]];
    c.K *= dims[axes[m]];
  }
  c.nrest = 0;
  uint64_t groups = 1;
  for (int i = 0; i < rank; ++i)
    if (!is_gate[i]) {
      c.rest_dims[c.nrest] = dims[i];
      c.rest_st[c.nrest] = st[i];
      ++c.nrest;
      groups *= dims[i];
    }
  c.state = state;
  c.G = G;
  c.k = k;
  parallel_range(groups, nthreads, gate_range, &c);
  return OR_OK;
}

struct diag_ctx {
  double *state;
  const double *d;
  int k;
  const uint64_t *dims;
  const uint16_t *axes;
  uint64_t st[MAXR], dst[MAXR];
};

static void diag_range(void *p, uint64_t b, uint64_t e) {
  const struct diag_ctx *c = (const struct diag_ctx *)p;
  for (uint64_t x = b; x < e; ++x) {
    uint64_t di = 0;
    for (int m = 0; m < c->k; ++m) di += ((x / c->st[c->axes[m]]) % c->dims[c->axes[m]]) * c->dst[m];
    double ar = c->state[2 * x], ai = c->state[2 * x + 1];
    double br = c->d[2 * di], bi = c->d[2 * di + 1];
    c->state[2 * x] = ar * br - ai * bi;
    c->state[2 * x + 1] = ar * bi + ai * br;
  }
}

int or_diag_apply_mt(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes, const double *d,
                     int nthreads) {
  if (k <= 0 || k > rank || rank > MAXR) return OR_EINVAL;
  struct diag_ctx c;
  strides_of(rank, dims, c.st);
  uint64_t s = 1;
  for (int m = k - 1; m >= 0; --m) {
    if (axes[m] >= rank) return OR_EINVAL;
    c.dst[m] = s;
    s *= dims[axes[m]];
  }
  c.state = state;
  c.d = d;
  c.k = k;
  c.dims = dims;
  c.axes = axes;
  parallel_range(or_size(rank, dims), nthreads, diag_range, &c);
  return OR_OK;
}
