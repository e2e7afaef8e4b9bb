/*
 * ctas_oracle.c - CPU restatement of the CTAS filtered-backprojection path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg. The engine
 * never links it. Build with -ffp-contract=off and without -ffast-math (see
 * oracle/Makefile) so that every float product and sum is rounded separately,
 * as in an x86-64 build of the reference; liboracle_fast.so is the same code
 * built with the reference's own -O3 -ffast-math flags, used only for timing.
 *
 * Reference citations are file:line in the CTAS source tree.
 */
#define _GNU_SOURCE
#include "ctas_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* geometry helpers                                                    */
/* ------------------------------------------------------------------ */

/* lct.cpp:490 - zidth = 2^ceil(log2(2w-1)): the smallest power of two that
 * is >= 2w-1. */
int orc_zidth(int w) {
  if (w < 1) return 0;
  int z = 1;
  while (z < 2 * w - 1) z <<= 1;
  return z;
}

/* lct.cpp:470-476 */
float orc_arcO(float anarc, int points) {
  float r = anarc;
  if (r < 1.0f) r *= (float)(points - 1);
  r = (float)((double)r * (M_PI / 180));
  return r;
}

/* lct.cpp:478-484 */
float orc_angO(float anarc, int points) {
  float r = anarc;
  if (r > 1.0f && points > 1) r /= (float)(points - 1);
  r = (float)((double)r * (M_PI / 180));
  return r;
}

/* lct.cpp:512-517 - th = i*arc/N in float, table entry {.x = sinf, .y = cosf}. */
void orc_angles(int n, float anarc, float *sincos) {
  const float arc = orc_arcO(anarc, n);
  for (int i = 0; i < n; i++) {
    float th = (float)i * arc;
    th = th / (float)n;
    sincos[2 * i] = sinf(th);
    sincos[2 * i + 1] = cosf(th);
  }
}

/* ------------------------------------------------------------------ */
/* FP64 radix-2 complex FFT (in place, forward sign -1, unnormalised)  */
/* ------------------------------------------------------------------ */

static void fft64(double *re, double *im, int n, int inverse) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const double ang = (inverse ? 2.0 : -2.0) * M_PI / len;
    const int half = len >> 1;
    for (int k = 0; k < half; k++) {
      const double wr = cos(ang * k), wi = sin(ang * k);
      for (int i = k; i < n; i += len) {
        const int j = i + half;
        const double xr = re[j] * wr - im[j] * wi;
        const double xi = re[j] * wi + im[j] * wr;
        re[j] = re[i] - xr; im[j] = im[i] - xi;
        re[i] += xr; im[i] += xi;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* filter window                                                       */
/* ------------------------------------------------------------------ */

/* Modified Bessel I0 by its power series (GSL's gsl_sf_bessel_I0 is
 * accurate to double precision; so is this series for the alphas used). */
static double bessel_i0(double x) {
  double sum = 1.0, term = 1.0;
  const double q = 0.25 * x * x;
  for (int k = 1; k < 500; k++) {
    term *= q / ((double)k * k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

/* Filter::fill (lct.cpp:122-209) for a line of n points. The spatial
 * Ram-Lak taps h[0] = 1/4, h[m] = -1/(m pi)^2 (odd m < n/2, integer
 * division) are built in double and stored as float (lct.cpp:132-135);
 * FFTW's HC2R of them, H[k] = h[0] + 2 sum h[m] cos(2 pi m k / n) (a zero
 * imaginary half), is evaluated here in FP64 (an FFT for a power of two n,
 * the direct sum otherwise) and rounded once to float. Each value is then
 * multiplied by the window of f = 2 min(k, n - k) / n, halved. */
int orc_filter_fill(int ftype, float param, int n, float *W) {
  if (n <= 0) return -1;
  if (ftype == ORC_KAISER && param < 0.0f) return -2;
  if (ftype == ORC_GAUSS && param <= 0.0f) return -2;
  double *re = (double *)calloc((size_t)n, sizeof(double));
  double *im = (double *)calloc((size_t)n, sizeof(double));
  if (!re || !im) { free(re); free(im); return -3; }
  re[0] = (double)0.25f;
  for (int m = 1; m < n / 2; m += 2) {
    const float h = (float)(-1.0 / ((double)(m * m) * M_PI * M_PI));
    re[m] = (double)h;
    re[n - m] = (double)h;
  }
  if ((n & (n - 1)) == 0) {
    fft64(re, im, n, 0);
  } else {
    /* the even extension above makes the sum over all n entries equal the
     * halfcomplex HC2R sum */
    for (int k = 0; k < n; k++) {
      double acc = 0.0;
      for (int m = 0; m < n; m++)
        if (re[m] != 0.0) acc += re[m] * cos(2.0 * M_PI * (double)(((long long)m * k) % n) / n);
      im[k] = acc;
    }
    memcpy(re, im, sizeof(double) * (size_t)n);
  }
  for (int k = 0; k < n; k++) {
    float filt = (float)re[k];
    float fp = (float)((k < n / 2) ? k : n - k);
    fp = (float)((double)fp * (2.0 / n));
    switch (ftype) {
    case ORC_NONE:
    case ORC_RAMP: fp = 1.0f; break;
    case ORC_BARLETT: fp = 1 - fp; break;
    case ORC_WELCH: fp = 1 - fp * fp; break;
    case ORC_PARZEN:
      fp = (fp <= 0.5f) ? 1 - 6 * fp * fp + 3 * fp * fp * fp : 2 * (1 + fp * fp * fp);
      break;
    case ORC_HANN: fp = (float)(0.5 * (1 + cos(M_PI * fp))); break;
    case ORC_HAMMING: fp = (float)(0.54 + 0.46 * cos(M_PI * fp)); break;
    case ORC_BLACKMAN:
      fp = (float)(0.42 + 0.5 * cos(M_PI * fp) + 0.08 * cos(2 * M_PI * fp));
      break;
    case ORC_LANCKZOS: fp = (fp == 0.0f) ? 0.0f : (float)(fabs(sin(M_PI * fp)) / fp); break;
    case ORC_KAISER: {
      const float arg = param * sqrtf(1 - fp * fp);
      fp = (float)(bessel_i0((double)arg) / bessel_i0((double)param));
      break;
    }
    case ORC_GAUSS: fp = powf(2.0f, (-fp) * fp / (param * param)); break;
    default: free(re); free(im); return -2;
    }
    filt = (float)((double)filt * ((double)fp / 2.0));
    W[k] = filt;
  }
  free(re);
  free(im);
  return 0;
}

/* CTrec's window (lct.cpp:487-508): Filter::fill at the padded width Z,
 * times the gain Z/w; REF: no DC. */
int orc_filter_window(int ftype, float param, int Z, int w, int contrast, float *W) {
  if (Z <= 0 || w <= 0) return -1;
  const int rc = orc_filter_fill(ftype, param, Z, W);
  if (rc) return rc;
  const float gain = (float)Z / (float)w;
  for (int k = 0; k < Z; k++) W[k] = W[k] * gain;
  if (contrast == ORC_REF) W[0] = 0.0f;
  return 0;
}

/* ------------------------------------------------------------------ */
/* deAbs                                                               */
/* ------------------------------------------------------------------ */

/* physical.world.cpp:88-100 */
int orc_deabs(float *a, long n) {
  if (n <= 0) return 0;
  float mn = a[0], mx = a[0];
  for (long i = 1; i < n; i++) {
    if (a[i] < mn) mn = a[i];
    if (a[i] > mx) mx = a[i];
  }
  if (mn <= 0.0f) {
    const float mina = (float)((double)mx / 1000000.0);
    if (mina <= 0.0f) return -1;
    for (long i = 0; i < n; i++)
      if (a[i] <= mina) a[i] = mina;
  }
  for (long i = 0; i < n; i++) a[i] = -logf(a[i]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* row filtering                                                       */
/* ------------------------------------------------------------------ */

/* prepare_sino's filtering loop (lct.cpp:577-588) with filter_line
 * (lct.cpp:280-291). Row theta is copied into the zeroed padded line at
 * s0 = (Z-w)/2 with the inclusive blitz range, so z[s0+w] receives the
 * first sample of row theta+1 (0 for the last row, where the reference
 * reads past the array). The line is convolved with IDFT(W) and the
 * centre w samples are written back. Rows are processed in pairs as the
 * real and imaginary parts of one FP64 complex FFT: W is real and even,
 * so the two filtered rows separate exactly. */
void orc_filter_rows(float *P, int N, int w, const float *W, int Z) {
  const int s0 = (Z - w) / 2;
  double *re = (double *)malloc(sizeof(double) * (size_t)Z);
  double *im = (double *)malloc(sizeof(double) * (size_t)Z);
  for (int t = 0; t < N; t += 2) {
    const int has2 = (t + 1 < N);
    memset(re, 0, sizeof(double) * (size_t)Z);
    memset(im, 0, sizeof(double) * (size_t)Z);
    for (int x = 0; x < w; x++) re[s0 + x] = P[(size_t)t * w + x];
    re[s0 + w] = (t + 1 < N) ? P[(size_t)(t + 1) * w] : 0.0;
    if (has2) {
      for (int x = 0; x < w; x++) im[s0 + x] = P[(size_t)(t + 1) * w + x];
      im[s0 + w] = (t + 2 < N) ? P[(size_t)(t + 2) * w] : 0.0;
    }
    fft64(re, im, Z, 0);
    for (int k = 0; k < Z; k++) {
      re[k] *= (double)W[k];
      im[k] *= (double)W[k];
    }
    fft64(re, im, Z, 1);
    for (int x = 0; x < w; x++) P[(size_t)t * w + x] = (float)(re[s0 + x] / Z);
    if (has2)
      for (int x = 0; x < w; x++) P[(size_t)(t + 1) * w + x] = (float)(im[s0 + x] / Z);
  }
  free(re);
  free(im);
}

/* CTrec::prepare_sino (lct.cpp:569-597). Returns <0 where the reference
 * throws (deAbs on data whose maximum is not positive). */
int orc_prepare_sino(float *P, int N, int w, const float *W, int Z, int contrast) {
  if (contrast == ORC_ABS && orc_deabs(P, (long)N * w) < 0) return -1;
  if (contrast != ORC_FLT) orc_filter_rows(P, N, w, W, Z);
  if (contrast == ORC_REF) {
    for (int t = 0; t < N; t++) {
      float *row = P + (size_t)t * w;
      for (int x = 1; x < w; x++) row[x] = row[x - 1] + row[x];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* backprojection                                                      */
/* ------------------------------------------------------------------ */

/* CTrec::reconstructOnCPU (lct.cpp:688-729) and ct.cl:5-44, in float with
 * the reference's evaluation order: ((c + (1-s-co)*hp) + s*(j-c)) + co*i,
 * truncated to an integer. The sample is read at flat offset theta*w + idx
 * of the slice's sinogram buffer, as the reference does; offsets outside
 * the buffer (never reached inside the mask) read 0. */
void orc_bp_ctas(const float *P, int N, int w, const float *sincos, float c,
                 float recCof, float *out) {
  orc_bp_ctas_rows(P, N, w, sincos, c, recCof, NULL, w, out);
}

/* orc_bp_ctas for the slice rows rows[0..n_rows) only (all rows when rows is
 * NULL): out[k * w + i] is pixel (rows[k], i), the same float operations. */
void orc_bp_ctas_rows(const float *P, int N, int w, const float *sincos, float c,
                      float recCof, const int *rows, int n_rows, float *out) {
  const int hp = w / 2;
  const float acent = fabsf(c);
  const long total_len = (long)N * w;
  for (long k = 0; k < n_rows; k++) {
    const long j = rows ? rows[k] : k;
    for (long i = 0; i < w; i++) {
      int ip2 = (int)i - hp;
      ip2 *= ip2;
      if ((float)i <= acent || (float)i >= (float)w - acent - 1) {
        out[k * w + i] = 0.0f;
        continue;
      }
      float ijp2 = (float)(j - hp) - c;
      ijp2 *= ijp2;
      ijp2 += (float)ip2;
      const float up = (float)hp + c, dn = (float)hp - c;
      if (((float)j <= up && ijp2 >= up * up) || ((float)j >= up && ijp2 >= dn * dn)) {
        out[k * w + i] = 0.0f;
        continue;
      }
      float acc = 0.0f;
      for (long p = 0; p < N; p++) {
        const float s = sincos[2 * p], co = sincos[2 * p + 1];
        float t = c + ((1.0f - s) - co) * (float)hp;
        t = t + s * ((float)j - c);
        t = t + co * (float)i;
        const long flat = p * w + (long)t;
        acc += (flat >= 0 && flat < total_len) ? P[flat] : 0.0f;
      }
      if (recCof != 0.0f) acc *= recCof;
      out[k * w + i] = acc / (float)N;
    }
  }
}

/* sin and cos of th in FP64, exact (0, +-1) at multiples of pi/2, where the
 * ray coordinate of many pixels is an exact integer and a residue such as
 * cos(pi/2) = 6e-17 would decide the floor. */
static void sincos_exact(double th, double *s, double *c) {
  const double q = th / (M_PI / 2);
  const double k = floor(q + 0.5);
  if (fabs(q - k) < 1e-12) {
    const long m = ((long)k % 4 + 4) % 4;
    *s = (m == 1) ? 1.0 : (m == 3) ? -1.0 : 0.0;
    *c = (m == 0) ? 1.0 : (m == 2) ? -1.0 : 0.0;
    return;
  }
  *s = sin(th);
  *c = cos(th);
}

/* FP64 oracle of the same discretisation: angles i*arc/N and the ray
 * coordinate in double, same masks. linear == 0: the reference's sampling
 * (lct.cpp:710-713), the sample at the flat offset p*w + (long)t of the
 * sinogram (truncation; a coordinate outside the row reads the neighbouring
 * row, 0 outside the sinogram). linear == 1: linear interpolation between
 * bin centres k+1/2, zero outside the detector. */
void orc_bp_f64(const float *P, int N, int w, double arc_rad, double c,
                double recCof, int linear, double *out) {
  orc_bp_f64_rows(P, N, w, arc_rad, c, recCof, linear, NULL, w, out);
}

/* orc_bp_f64 for the slice rows rows[0..n_rows) only (all rows when rows is
 * NULL): out[k * w + i] is pixel (rows[k], i). */
void orc_bp_f64_rows(const float *P, int N, int w, double arc_rad, double c,
                     double recCof, int linear, const int *rows, int n_rows, double *out) {
  const int hp = w / 2;
  const double acent = fabs(c);
  const long total_len = (long)N * w;
  double *S = (double *)malloc(sizeof(double) * (size_t)N);
  double *C = (double *)malloc(sizeof(double) * (size_t)N);
  for (int p = 0; p < N; p++) sincos_exact(p * arc_rad / N, &S[p], &C[p]);
  for (long k = 0; k < n_rows; k++) {
    const long j = rows ? rows[k] : k;
    for (long i = 0; i < w; i++) {
      const double dy = (double)(j - hp) - c;
      const double d2 = dy * dy + (double)((i - hp) * (i - hp));
      const double up = hp + c, dn = hp - c;
      if (i <= acent || i >= w - acent - 1 || (j <= up && d2 >= up * up) ||
          (j >= up && d2 >= dn * dn)) {
        out[k * w + i] = 0.0;
        continue;
      }
      double acc = 0.0;
      for (int p = 0; p < N; p++) {
        const double t = c + hp + S[p] * (j - c - hp) + C[p] * (i - hp);
        const float *row = P + (size_t)p * w;
        if (!linear) {
          const long flat = (long)p * w + (long)t;
          acc += (flat >= 0 && flat < total_len) ? P[flat] : 0.0;
        } else {
          const double u = t - 0.5;
          const long k = (long)floor(u);
          const double f = u - k;
          const double a = (k >= 0 && k < w) ? row[k] : 0.0;
          const double b = (k + 1 >= 0 && k + 1 < w) ? row[k + 1] : 0.0;
          acc += a + f * (b - a);
        }
      }
      if (recCof != 0.0) acc *= recCof;
      out[k * w + i] = acc / N;
    }
  }
  free(S);
  free(C);
}

/* ------------------------------------------------------------------ */
/* threaded batch: the RecInThread model (ct.cpp:156-236)              */
/* ------------------------------------------------------------------ */

typedef struct {
  float *sino, *out;
  const float *W, *sincos, *centers;
  int S, N, w, Z, contrast;
  float recCof;
  int next;
  int err;
  pthread_mutex_t mu;
} batch_ctx;

static void *batch_worker(void *arg) {
  batch_ctx *b = (batch_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(&b->mu);
    const int s = b->next++;
    pthread_mutex_unlock(&b->mu);
    if (s >= b->S) break;
    float *P = b->sino + (size_t)s * b->N * b->w;
    if (orc_prepare_sino(P, b->N, b->w, b->W, b->Z, b->contrast) < 0) {
      pthread_mutex_lock(&b->mu);
      b->err = -1;
      pthread_mutex_unlock(&b->mu);
      continue;
    }
    orc_bp_ctas(P, b->N, b->w, b->sincos, b->centers[s], b->recCof,
                b->out + (size_t)s * b->w * b->w);
  }
  return NULL;
}

/* S slices: prepare_sino + reconstructOnCPU, one slice per thread. The
 * sinograms are filtered in place, as CTrec::reconstruct does. */
int orc_fbp_batch(float *sino, int S, int N, int w, int contrast, int ftype,
                  float fparam, float anarc, const float *centers, float recCof,
                  float *out, int nthreads) {
  const int Z = orc_zidth(w);
  float *W = (float *)malloc(sizeof(float) * (size_t)Z);
  float *sc = (float *)malloc(sizeof(float) * 2 * (size_t)N);
  if (orc_filter_window(ftype, fparam, Z, w, contrast, W) < 0) {
    free(W); free(sc);
    return -2;
  }
  orc_angles(N, anarc, sc);
  batch_ctx b;
  b.sino = sino; b.out = out; b.W = W; b.sincos = sc; b.centers = centers;
  b.S = S; b.N = N; b.w = w; b.Z = Z; b.contrast = contrast; b.recCof = recCof;
  b.next = 0; b.err = 0;
  pthread_mutex_init(&b.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, batch_worker, &b);
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&b.mu);
  free(th);
  free(W);
  free(sc);
  return b.err;
}

/* ------------------------------------------------------------------ */
/* ring filter                                                         */
/* ------------------------------------------------------------------ */

/* RingFilter::apply (lct.cpp:295-327) without a mask (the ct tool passes
 * none, ct.cpp:200). */
int orc_ring_filter(float *P, int N, int w, int box) {
  if (box <= 0 || N <= 0) return 0;
  /* temp(ii) and temp(w-1-ii) for every ii < box (lct.cpp:317-320) lie in
     the line only when box < w; wider boxes index outside it there */
  if (box >= w) return -1;
  float *avg = (float *)calloc((size_t)w, sizeof(float));
  float *tmp = (float *)calloc((size_t)w, sizeof(float));
  for (int r = 0; r < N; r++)
    for (int x = 0; x < w; x++) avg[x] = avg[x] + P[(size_t)r * w + x];
  for (int x = 0; x < w; x++) avg[x] = avg[x] / (float)N;
  for (int ii = -box; ii <= box; ii++) {
    const int d0 = ii > 0 ? ii : 0;
    const int d1 = w - 1 + (ii < 0 ? ii : 0);
    for (int x = d0; x <= d1; x++) tmp[x] = tmp[x] + avg[x - ii];
  }
  for (int x = box; x <= w - 1 - box; x++) tmp[x] = tmp[x] / (float)(2 * box + 1);
  for (int ii = 0; ii < box; ii++) {
    tmp[ii] = tmp[ii] / (float)(box + ii + 1);
    tmp[w - 1 - ii] = tmp[w - 1 - ii] / (float)(box + ii + 1);
  }
  for (int x = 0; x < w; x++) tmp[x] = tmp[x] - avg[x];
  for (int r = 0; r < N; r++)
    for (int x = 0; x < w; x++) P[(size_t)r * w + x] = P[(size_t)r * w + x] + tmp[x];
  free(avg);
  free(tmp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* axis finding                                                        */
/* ------------------------------------------------------------------ */

/* AverageGradient::apply (lct.cpp:872-915) on an already prepared (ABS,
 * RAMP) sinogram. The one-sided Gaussian smoothing of lct.cpp:837-843 and
 * :878-884 is the circular convolution with g[u]*g[v], g[n] = exp(-n^2 /
 * sig22); it is evaluated here separably in FP64 (the FFT of the reference
 * computes the same sum in float). The gradient magnitude is read from the
 * top rows of the full derivative arrays, reproducing the shape mismatch of
 * lct.cpp:905. work may be NULL. Parity note: the reference's positive-value
 * count depends on FFTW round-off; the restatement counts every pixel of the
 * derivative rows when smoothing (see the count below). */
double orc_ag_score(const float *Pf, int N, int w, const float *sincos,
                    float sigma, float axisR, float *work) {
  (void)work;
  const float centre = (float)((double)axisR - 0.5 * (w - 1));
  const size_t ww = (size_t)w * w;
  float *sl = (float *)malloc(sizeof(float) * ww);
  orc_bp_ctas(Pf, N, w, sincos, centre, 1.0f, sl);
  if (sigma > 1) {
    const float sig22 = 2 * sigma * sigma;
    double *g = (double *)malloc(sizeof(double) * (size_t)w);
    int nt = 0;
    for (int n = 0; n < w; n++) {
      g[n] = exp(-((double)n * n) / (double)sig22);
      if (g[n] > 1e-18) nt = n + 1;
    }
    double *tmp = (double *)malloc(sizeof(double) * ww);
    for (int u = 0; u < w; u++)
      for (int v = 0; v < w; v++) {
        double a = 0;
        for (int n = 0; n < nt; n++) a += g[n] * sl[(size_t)u * w + ((v - n) % w + w) % w];
        tmp[(size_t)u * w + v] = a;
      }
    for (int u = 0; u < w; u++)
      for (int v = 0; v < w; v++) {
        double a = 0;
        for (int m = 0; m < nt; m++) a += g[m] * tmp[(size_t)(((u - m) % w + w) % w) * w + v];
        sl[(size_t)u * w + v] = (float)a;
      }
    free(tmp);
    free(g);
  }
  const int r = (int)ceilf(fabsf(centre));
  const int h = w - 2 * r;
  if (h < 3) { free(sl); return 0.0; }
  const float *cs = sl + (size_t)r * w; /* cropped view, h x w */
  float *dH = (float *)calloc(ww, sizeof(float));
  float *dV = (float *)calloc(ww, sizeof(float));
  static const int hs[6][3] = {{-1, -1, -1}, {-1, 0, -2}, {-1, 1, -1},
                               {1, -1, 1},   {1, 0, 2},   {1, 1, 1}};
  static const int vs[6][3] = {{-1, -1, -1}, {0, -1, -2}, {1, -1, -1},
                               {-1, 1, 1},   {0, 1, 2},   {1, 1, 1}};
  for (int q = 0; q < 6; q++) {
    for (int y = 0; y < h; y++) {
      const int sy = y - hs[q][0];
      if (sy < 0 || sy >= h) continue;
      for (int x = 0; x < w; x++) {
        const int sx = x - hs[q][1];
        if (sx < 0 || sx >= w) continue;
        float *d = dH + (size_t)(y + r) * w + x;
        *d = *d + (float)hs[q][2] * cs[(size_t)sy * w + sx];
      }
    }
  }
  for (int q = 0; q < 6; q++) {
    for (int y = 0; y < h; y++) {
      const int sy = y - vs[q][0];
      if (sy < 0 || sy >= h) continue;
      for (int x = 0; x < w; x++) {
        const int sx = x - vs[q][1];
        if (sx < 0 || sx >= w) continue;
        float *d = dV + (size_t)(y + r) * w + x;
        *d = *d + (float)vs[q][2] * cs[(size_t)sy * w + sx];
      }
    }
  }
  /* The count #(cS > 0) of lct.cpp:912: after the reference's float FFT
   * smoothing every value carries round-off noise, so every pixel of the rows
   * that received derivatives (full-array rows >= r) counts; without
   * smoothing, exact zeros (the masked corners) do not. */
  const int smoothed = sigma > 1;
  double sm = 0.0;
  long nm = 0;
  for (int y = 1; y < h - 1; y++) {
    for (int x = 1; x < w - 1; x++) {
      const float a = dH[(size_t)y * w + x], b = dV[(size_t)y * w + x];
      const float m = sqrtf(a * a + b * b);
      sm += m;
      nm += smoothed ? (y >= r) : (m > 0.0f);
    }
  }
  free(dH);
  free(dV);
  free(sl);
  return nm ? sm / (double)nm : 0.0;
}

/* improveMe (lct.cpp:1108-1129) */
typedef double (*score_fn)(void *ctx, float x);

static float improve_me(float eps, const float *vax, const float *err, int idx,
                        int cdir, score_fn f, void *ctx) {
  float aa = vax[idx - cdir];
  float bb = vax[idx + cdir];
  float rf = vax[idx];
  float vl = err[idx];
  while (fabsf(aa - rf) > eps || fabsf(rf - bb) > eps) {
    const int afar = fabsf(aa - rf) > fabsf(rf - bb);
    float *lo = afar ? &aa : &bb;
    float *hi = afar ? &bb : &aa;
    const float x = (rf + *lo) / 2;
    const float vx = (float)f(ctx, x);
    if (vl <= vx) {
      *lo = x;
    } else {
      vl = vx;
      *hi = rf;
      rf = x;
    }
  }
  return rf;
}

typedef struct {
  const float *Pf, *sincos;
  int N, w;
  float sigma, cent;
} ag_ctx;

static double ag_eval(void *vctx, float x) {
  ag_ctx *c = (ag_ctx *)vctx;
  return (float)orc_ag_score(c->Pf, c->N, c->w, c->sincos, c->sigma, c->cent + x, NULL);
}

typedef struct {
  const float *p0, *p180r;
  int rows, w;
  float aver;
} pair_ctx;

/* calcErr of raxis(proj0, proj180, maxDev) (lct.cpp:1264-1271). */
static double pair_err(void *vctx, float x) {
  pair_ctx *c = (pair_ctx *)vctx;
  const long sh = (long)x; /* dstR/srcR take an integer shift */
  const int w = c->w;
  double acc = 0.0;
  for (int r = 0; r < c->rows; r++) {
    for (int col = 0; col < w; col++) {
      float d = c->aver;
      const long src = col - sh;
      if (col >= (sh > 0 ? sh : 0) && col <= w - 1 + (sh < 0 ? sh : 0) && src >= 0 && src < w)
        d = c->p180r[(size_t)r * w + src];
      d = d - c->p0[(size_t)r * w + col];
      acc += (double)(d * d);
    }
  }
  return sqrtf((float)acc);
}

/* raxis(proj0, proj180, maxDev) (lct.cpp:1217-1294). */
int orc_raxis_pair(const float *p0, const float *p180, int rows, int w, float maxDev,
                   float *result) {
  if (maxDev < 0 || maxDev > 0.5f) return -1;
  if (rows <= 0 || w <= 0) return -2;
  const float cent = (float)(0.5 * (w - 1));
  float *proj = (float *)calloc((size_t)w, sizeof(float));
  for (int r = 0; r < rows; r++)
    for (int x = 0; x < w; x++)
      proj[x] = (float)((double)proj[x] + (fabs((double)p0[(size_t)r * w + x] - 1.0) +
                                           fabs((double)p180[(size_t)r * w + x] - 1.0)));
  double ssum = 0.0, wsum = 0.0;
  for (int x = 0; x < w; x++) {
    ssum += proj[x];
    wsum += (double)((float)x * proj[x]);
  }
  free(proj);
  if (ssum == 0.0) return -3;
  const float fcr = (float)((float)wsum / (float)ssum);
  const int stpsz = 16;
  const int steps = (int)(1 + 2 * (maxDev == 0.0f ? 1.0f : maxDev * (float)w / stpsz));
  float *p180r = (float *)malloc(sizeof(float) * (size_t)rows * w);
  for (int r = 0; r < rows; r++)
    for (int x = 0; x < w; x++) p180r[(size_t)r * w + x] = p180[(size_t)r * w + (w - 1 - x)];
  double a0 = 0.0, a1 = 0.0;
  for (int r = 0; r < rows; r++) {
    a0 += p180r[(size_t)r * w];
    a1 += p180r[(size_t)r * w + w - 1];
  }
  pair_ctx ctx = {p0, p180r, rows, w, (float)(((float)a0 + (float)a1) / (float)(2 * rows))};
  float *vax = (float *)malloc(sizeof(float) * (size_t)(steps + 2));
  float *err = (float *)malloc(sizeof(float) * (size_t)(steps + 2));
  for (int ii = 0; ii < steps; ii++) {
    vax[ii] = 2 * (fcr - cent) + (float)(stpsz * (ii - (steps - 1) / 2));
    err[ii] = (float)pair_err(&ctx, vax[ii]);
  }
  int idx = -1, rc = 0;
  float minval = FLT_MAX;
  for (int i = 1; i < steps - 1; i++)
    if (err[i] < err[i - 1] && err[i] < err[i + 1] && err[i] < minval) minval = err[idx = i];
  if (idx >= 0) {
    *result = cent + 0.5f * improve_me(1.0f, vax, err, idx, 1, pair_err, &ctx);
    goto done;
  }
  idx = (err[0] < err[steps - 1]) ? 0 : steps - 1;
  {
    const int cdir = (idx == 0) ? -1 : 1;
    minval = err[idx];
    const int maxSteps = (int)(((float)(w / 2) - (float)cdir * vax[idx]) / stpsz);
    if (maxSteps <= 0) { rc = -4; goto done; }
    float *v2 = (float *)malloc(sizeof(float) * (size_t)(maxSteps + 2));
    float *e2 = (float *)malloc(sizeof(float) * (size_t)(maxSteps + 2));
    v2[0] = vax[idx - cdir]; v2[1] = vax[idx];
    e2[0] = err[idx - cdir]; e2[1] = err[idx];
    int idx2 = 1;
    for (int i = 0; i < maxSteps; i++) {
      v2[i + 2] = vax[idx] + (float)(cdir * (i + 1) * stpsz);
      e2[i + 2] = (float)pair_err(&ctx, v2[i + 2]);
      if (e2[i + 2] < minval) {
        minval = e2[i + 2];
        idx2 = i + 2;
      } else {
        break;
      }
    }
    if (idx2 == maxSteps + 1) rc = -4;
    else *result = cent + 0.5f * improve_me(1.0f, v2, e2, idx2, cdir, pair_err, &ctx);
    free(v2);
    free(e2);
  }
done:
  free(vax);
  free(err);
  free(p180r);
  return rc;
}

/* raxis(sino, anarc, maxDev, algo) (lct.cpp:1133-1212). The trial search is
 * evaluated sequentially; the directional walk stops at the first trial
 * that does not improve, the deterministic order of lct.cpp:1028-1040. */
int orc_raxis(const float *sino, int N, int w, float anarc, float maxDev, int algo,
              float *result) {
  if (N < 3 || w < 3) return -1;
  const float arc = orc_arcO(anarc, N);
  if (arc == 0.0f) return -2;
  if (fabsf(arc) < (float)M_PI) return -3;
  const float maxDevLM = 0.25f;
  if (maxDev < 0 || maxDev > maxDevLM) return -4;
  const float cent = (float)(0.5 * (w - 1));

  float fcr = 0;
  if (algo == 0 || algo == 2) {
    const int idx180 = (int)((N - 1) * M_PI / fabsf(orc_arcO(anarc, N)));
    const int nofln = N - idx180;
    if (nofln <= 0) return -5;
    float *r0 = (float *)malloc(sizeof(float) * (size_t)nofln * w);
    float *r180 = (float *)malloc(sizeof(float) * (size_t)nofln * w);
    for (int j = 0; j < nofln; j++)
      for (int x = 0; x < w; x++) {
        r0[(size_t)j * w + x] = sino[(size_t)j * w + x] + 1;
        r180[(size_t)j * w + x] = sino[(size_t)(j + idx180) * w + x] + 1;
      }
    float q = 0;
    const int rc = orc_raxis_pair(r0, r180, nofln, w, maxDev < 0.5f ? maxDev : 0.5f, &q);
    free(r0);
    free(r180);
    if (rc < 0) return rc;
    if (algo == 0) { *result = q; return 0; }
    fcr = q;
  } else if (algo == 1) {
    fcr = cent;
  } else if (algo == 3) {
    float *pl = (float *)calloc((size_t)w, sizeof(float));
    for (int j = 0; j < N; j++)
      for (int x = 0; x < w; x++) pl[x] = pl[x] + sino[(size_t)j * w + x];
    float mx = pl[0], mn = pl[0];
    for (int x = 1; x < w; x++) {
      if (pl[x] > mx) mx = pl[x];
      if (pl[x] < mn) mn = pl[x];
    }
    /* lct.cpp:1168-1180: the loop condition is always true, so the search
     * only ends at the first sample above the threshold; with a constant
     * projection it never ends in the reference. */
    if (!(mx > mn)) { free(pl); return -6; }
    const float thr = (float)(0.1 * mx + 0.9 * mn);
    int a = 0, b = w - 1;
    while (!(pl[a] > thr)) a++;
    while (!(pl[b] > thr)) b--;
    free(pl);
    fcr = ((float)a + (float)b) / 2.0f;
  } else {
    return -7;
  }

  if (fabsf(fcr - cent) / w > maxDev) return -8;
  const float rbxsz = (float)(1.0 - 2.0 * (maxDev + fabsf(fcr - cent) / w));
  if (rbxsz < 0.5f) return -9;

  /* AverageGradient ctor (lct.cpp:818-844): ABS contrast, default RAMP. */
  const int Z = orc_zidth(w);
  float *W = (float *)malloc(sizeof(float) * (size_t)Z);
  float *sc = (float *)malloc(sizeof(float) * 2 * (size_t)N);
  float *Pf = (float *)malloc(sizeof(float) * (size_t)N * w);
  memcpy(Pf, sino, sizeof(float) * (size_t)N * w);
  orc_filter_window(ORC_RAMP, 0.0f, Z, w, ORC_ABS, W);
  orc_angles(N, anarc, sc);
  int rc = orc_prepare_sino(Pf, N, w, W, Z, ORC_ABS);
  ag_ctx ctx = {Pf, sc, N, w, (float)(orc_angO(anarc, N) * (float)w / M_PI), cent};
  float *vax = NULL, *err = NULL;
  if (rc < 0) goto out;
  {
    const int stpsz = 8;
    const int steps = (int)(1 + 2 * (maxDev == 0.0f ? 1.0f : maxDev * (float)w / stpsz));
    const int hfsteps = (steps - 1) / 2;
    const float coffee = fcr - cent;
    vax = (float *)malloc(sizeof(float) * (size_t)(steps + 2));
    err = (float *)malloc(sizeof(float) * (size_t)(steps + 2));
    for (int k = 0; k < steps; k++) {
      vax[k] = coffee + (float)(stpsz * (k - hfsteps));
      err[k] = (float)ag_eval(&ctx, vax[k]);
    }
    int idxM = 0;
    float minval = FLT_MAX;
    for (int i = 1; i < steps - 1; i++)
      if (err[i] < err[i - 1] && err[i] < err[i + 1] && err[i] < minval) {
        idxM = i;
        minval = err[i];
      }
    int cdir = 1;
    if (!idxM) {
      const int idx = (err[0] < err[steps - 1]) ? 0 : steps - 1;
      cdir = (idx == 0) ? -1 : 1;
      const int coff2 = (int)vax[idx - cdir * 2];
      const int maxSteps = (int)((float)w * maxDevLM - (float)(cdir * coff2 / stpsz));
      if (maxSteps <= 0) { rc = -10; goto out; }
      free(vax);
      free(err);
      vax = (float *)malloc(sizeof(float) * (size_t)(maxSteps + 2));
      err = (float *)malloc(sizeof(float) * (size_t)(maxSteps + 2));
      for (int k = 0; k < maxSteps; k++) {
        vax[k] = (float)coff2 + (float)(stpsz * k * cdir);
        err[k] = (float)ag_eval(&ctx, vax[k]);
        if (err[k] < minval) {
          minval = err[k];
          idxM = k;
        } else {
          /* the walk stops here; vax[k] is kept for improveMe */
          break;
        }
      }
      if (idxM == maxSteps - 1 || idxM == 0) { *result = fcr; goto out; }
    }
    *result = cent + improve_me(0.25f, vax, err, idxM, cdir, ag_eval, &ctx);
  }
out:
  free(vax);
  free(err);
  free(W);
  free(sc);
  free(Pf);
  return rc;
}

/* ------------------------------------------------------------------ */
/* synthetic ellipse phantoms                                          */
/* ------------------------------------------------------------------ */

static unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* Line integrals of n_ell ellipses (x0, y0, a, b, phi, mu per px; pixel
 * units, x = i - hp, y = j - hp - c as in the backprojector) at detector bin
 * centres s = k + 1/2 - hp - c and angles theta = p*pi/N, in FP64. With
 * N0 > 0 the output is the noisy transmission counts/N0, counts drawn from
 * Poisson(N0 exp(-p)) (normal approximation above 64 expected counts);
 * otherwise it is the line integral p. */
void orc_phantom_sino(int N, int w, double c, int n_ell, const double *ell, double N0,
                      unsigned long long seed, float *out) {
  const int hp = w / 2;
  for (int p = 0; p < N; p++) {
    const double th = p * M_PI / N;
    const double ct = cos(th), st = sin(th);
    for (int k = 0; k < w; k++) {
      const double s = k + 0.5 - hp - c;
      double acc = 0.0;
      for (int e = 0; e < n_ell; e++) {
        const double *E = ell + 6 * e;
        const double s0 = E[0] * ct + E[1] * st;
        const double cp = cos(th - E[4]), sp = sin(th - E[4]);
        const double r2 = E[2] * E[2] * cp * cp + E[3] * E[3] * sp * sp;
        const double ds = s - s0;
        if (ds * ds < r2) acc += 2.0 * E[5] * E[2] * E[3] * sqrt(r2 - ds * ds) / r2;
      }
      double v = acc;
      if (N0 > 0) {
        const double lam = N0 * exp(-acc);
        const unsigned long long h1 = splitmix64(seed ^ splitmix64((unsigned long long)p * w + k));
        const unsigned long long h2 = splitmix64(h1);
        const double u1 = ((h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        double cnt;
        if (lam < 64.0) {
          const double L = exp(-lam);
          double prod = u1;
          unsigned long long st2 = h2;
          cnt = 0;
          while (prod > L && cnt < 1000) {
            st2 = splitmix64(st2);
            prod *= ((st2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            cnt += 1;
          }
        } else {
          const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
          cnt = floor(lam + sqrt(lam) * z + 0.5);
          if (cnt < 0) cnt = 0;
        }
        v = cnt / N0;
      }
      out[(size_t)p * w + k] = (float)v;
    }
  }
}

/* ------------------------------------------------------------------ */
/* tomosynthesis                                                       */
/* ------------------------------------------------------------------ */

/* filter_line (lct.cpp:280-291) on one row of n samples, by the defining
 * sums of FFTW's transforms in FP64: R2HC gives hc[k] = sum x[m] cos(2 pi
 * k m / n) for k <= n/2 and hc[n-k] = -sum x[m] sin(2 pi k m / n) for
 * 0 < k < n - k; hc is multiplied by W position by position; HC2R gives
 * y[m] = hc[0] + 2 sum_{0<k<n-k} (hc[k] cos - hc[n-k] sin)(2 pi k m / n)
 * (+ hc[n/2] (-1)^m for even n); the row becomes y/n, rounded to float.
 * cs and sn hold cos and sin of 2 pi j / n, j < n. */
static void ts_filter_row(float *row, int n, const float *W, const double *cs, const double *sn,
                          double *hc, double *y) {
  for (int k = 0; k <= n / 2; k++) {
    double r = 0.0, i = 0.0;
    for (int m = 0; m < n; m++) {
      const long long j = ((long long)k * m) % n;
      r += (double)row[m] * cs[j];
      i -= (double)row[m] * sn[j];
    }
    hc[k] = r;
    if (k > 0 && k < n - k) hc[n - k] = i;
  }
  for (int k = 0; k < n; k++) hc[k] *= (double)W[k];
  for (int m = 0; m < n; m++) {
    double acc = hc[0];
    for (int k = 1; k < n - k; k++) {
      const long long j = ((long long)k * m) % n;
      acc += 2.0 * (hc[k] * cs[j] - hc[n - k] * sn[j]);
    }
    if (n % 2 == 0) acc += (m % 2 ? -1.0 : 1.0) * hc[n / 2];
    y[m] = acc;
  }
  for (int m = 0; m < n; m++) row[m] = (float)(y[m] / n);
}

/* ts_add (lct.cpp:739-787). proj (rows x n) is replaced by its filtered
 * rows, as the reference's in-place row view leaves it: -log of every
 * sample for ABS (lct.cpp:744-745), filter_line with W = Filter::fill(n)
 * (FP64 sums, float result), and std::partial_sum in float for REF. result
 * gets result(y, x) += proj(y, (long)d), with d evaluated in float in the
 * reference's order and clamped to [0, n - 1] (lct.cpp:762-783). */
int orc_ts_add(float *proj, float *result, int rows, int n, const float *W, float center, int contrast,
               float angle, int plane) {
  if (rows <= 0 || n <= 0) return -1;
  if (contrast == ORC_ABS)
    for (long i = 0; i < (long)rows * n; i++) proj[i] = -logf(proj[i]);
  double *cs = (double *)malloc(sizeof(double) * (size_t)n);
  double *sn = (double *)malloc(sizeof(double) * (size_t)n);
  double *hc = (double *)malloc(sizeof(double) * (size_t)n);
  double *y = (double *)malloc(sizeof(double) * (size_t)n);
  if (!cs || !sn || !hc || !y) { free(cs); free(sn); free(hc); free(y); return -3; }
  for (int j = 0; j < n; j++) {
    cs[j] = cos(2.0 * M_PI * j / n);
    sn[j] = sin(2.0 * M_PI * j / n);
  }
  const float cur_sin = sinf(angle);
  const float cur_cos = cosf(angle);
  const float plane_cos = ((float)plane - center) * cur_sin;
  const float Rcenter = (float)(n * 0.5 + (double)center);
  for (int yc = 0; yc < rows; yc++) {
    float *row = proj + (size_t)yc * n;
    ts_filter_row(row, n, W, cs, sn, hc, y);
    if (contrast == ORC_REF)
      for (int x = 1; x < n; x++) row[x] = row[x - 1] + row[x];
    for (int x = 0; x < n; x++) {
      float di = ((float)x - Rcenter) * cur_cos - plane_cos + Rcenter;
      di = (di < 0) ? 0 : (di >= (float)n) ? (float)(n - 1) : di;
      result[(size_t)yc * n + x] += row[(long)di];
    }
  }
  free(cs);
  free(sn);
  free(hc);
  free(y);
  return 0;
}

/* ------------------------------------------------------------------ */
/* denoiser                                                            */
/* ------------------------------------------------------------------ */

/* Denoiser (matrix.world.cpp:952-1037) without a mask, in the reference's
 * array form: for every offset (i, j) of the disc, taken in order (i outer,
 * j inner, from -rad), tarr(dst) += iom(src) over the overlapping ranges
 * dst = [max(0, i), size - 1 + min(0, i)], src = dst - (i, j)
 * (matrix.world.h:261-272), and swghts counts the offsets; tarr is scaled by
 * invert(swghts) = 1.0 / n (double, stored as float) and the threshold rule
 * of proc() picks between the mean and the pixel. */
int orc_denoise(float *a, int H, int W, int rad, float thr) {
  if (rad < 0) rad = -rad;
  if (H <= 0 || W <= 0 || rad == 0) return 0;
  const size_t n = (size_t)H * W;
  float *tarr = (float *)calloc(n, sizeof(float));
  float *sw = (float *)calloc(n, sizeof(float));
  if (!tarr || !sw) { free(tarr); free(sw); return -3; }
  const unsigned r2 = (unsigned)(rad * rad);
  for (int i = -rad; i <= rad; i++)
    for (int j = -rad; j <= rad; j++) {
      if ((unsigned)(i * i + j * j) > r2) continue;
      const int y0 = i > 0 ? i : 0, y1 = H - 1 + (i < 0 ? i : 0);
      const int x0 = j > 0 ? j : 0, x1 = W - 1 + (j < 0 ? j : 0);
      for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) {
          tarr[(size_t)y * W + x] += a[(size_t)(y - i) * W + (x - j)];
          sw[(size_t)y * W + x] += 1.0f;
        }
    }
  for (size_t k = 0; k < n; k++) {
    const float inv = sw[k] == 0.0f ? 0.0f : (float)(1.0 / sw[k]);
    const float t = tarr[k] * inv;
    float *v = a + k;
    if (thr == 0.0f) *v = t;
    else if (thr < 0.0f) {
      if (*v == 0.0f || (t != 0.0f && fabsf((*v - t) / t) > -thr)) *v = t;
    } else if (*v == 0.0f || fabsf(*v - t) > thr) *v = t;
  }
  free(tarr);
  free(sw);
  return 0;
}

