/*
 * ctas_oracle.h - CPU restatement of the CTAS filtered-backprojection path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker for the HIP engine in
 * ctas_amd/. Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of
 * bench.py may load it; the engine never links, calls or falls back to it.
 *
 * Every function restates a reference function (file:line, paths relative to
 * the CTAS source tree) or is an FP64 oracle of the same discretisation.
 * Arrays are row-major float32: sinograms [N][w], slices [w][w].
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_zidth, orc_arcO, orc_angO      closed forms / hand-computed tables
 *   orc_filter_window                  FP64 closed-form Ram-Lak spectrum, numpy FFT
 *   orc_deabs                          closed form of the clamp rule
 *   orc_filter_rows                    delta KAT (Z/(2w))*h[n], numpy FFT convolution
 *   orc_bp_ctas                        constant-sinogram KAT, single-angle index KAT,
 *                                      brute-force Python loop on tiny inputs
 *   orc_bp_f64                         analytic ellipse phantom -> textbook FBP
 *   orc_ring_filter                    numpy restatement on random data + invariants
 *   orc_ag_score / orc_raxis*          known-centre phantoms, mirror symmetry,
 *                                      constant image -> 0 (statistical pins only)
 *   orc_phantom_sino                   analytic chord lengths of circles
 * FFTW's exact float rounding and GCC -ffast-math code generation in the
 * reference cannot be reproduced without a reference build (unbuildable here:
 * FFTW3, GSL, poptmx, Magick++ are absent); those effects are ~1e-7 relative.
 */
#ifndef CTAS_ORACLE_H
#define CTAS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Contrast and filter codes follow the reference enums
 * (physical.world.h:22-27, ct.h:95-146). */
enum { ORC_ABS = 0, ORC_REF = 1, ORC_PHS = 2, ORC_FLT = 3 };
enum { ORC_NONE = 0, ORC_RAMP, ORC_BARLETT, ORC_WELCH, ORC_PARZEN, ORC_HANN,
       ORC_HAMMING, ORC_BLACKMAN, ORC_LANCKZOS, ORC_KAISER, ORC_GAUSS };

int   orc_zidth(int w);
float orc_arcO(float anarc, int points);
float orc_angO(float anarc, int points);
void  orc_angles(int n, float anarc, float *sincos);
int   orc_filter_fill(int ftype, float param, int n, float *W);
int   orc_filter_window(int ftype, float param, int Z, int w, int contrast, float *W);
int   orc_deabs(float *a, long n);
void  orc_filter_rows(float *P, int N, int w, const float *W, int Z);
int   orc_prepare_sino(float *P, int N, int w, const float *W, int Z, int contrast);
void  orc_bp_ctas(const float *P, int N, int w, const float *sincos, float c,
                  float recCof, float *out);
void  orc_bp_ctas_rows(const float *P, int N, int w, const float *sincos, float c,
                       float recCof, const int *rows, int n_rows, float *out);
void  orc_bp_f64(const float *P, int N, int w, double arc_rad, double c,
                 double recCof, int linear, double *out);
void  orc_bp_f64_rows(const float *P, int N, int w, double arc_rad, double c,
                      double recCof, int linear, const int *rows, int n_rows,
                      double *out);
int   orc_fbp_batch(float *sino, int S, int N, int w, int contrast, int ftype,
                    float fparam, float anarc, const float *centers,
                    float recCof, float *out, int nthreads);
int   orc_ring_filter(float *P, int N, int w, int box);  /* -1: box >= w */
double orc_ag_score(const float *Pf, int N, int w, const float *sincos,
                    float sigma, float axisR, float *work);
int   orc_raxis(const float *sino, int N, int w, float anarc, float maxDev,
                int algo, float *result);
int   orc_raxis_pair(const float *p0, const float *p180, int rows, int w,
                     float maxDev, float *result);
int   orc_ts_add(float *proj, float *result, int rows, int n, const float *W,
                 float center, int contrast, float angle, int plane);
int   orc_denoise(float *a, int H, int W, int rad, float thr);
void  orc_phantom_sino(int N, int w, double c, int n_ell, const double *ell,
                       double N0, unsigned long long seed, float *out);

#ifdef __cplusplus
}
#endif
#endif
