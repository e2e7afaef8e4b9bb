// Host-side geometry of the reconstruction: padding width, arc and angle
// table, and the frequency window of each filter.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ctas_internal.hpp"

namespace ctas {

// lct.cpp:490 - zidth = 2^ceil(log2(2w-1)), the smallest power of two that
// holds a line and its zero padding.
int zidth(int w) {
  if (w < 1) return 0;
  int z = 1;
  while (z < 2 * w - 1) z <<= 1;
  return z;
}

// arcO (lct.cpp:470-476): an arc below one degree is a step, scaled by the
// number of intervals. The product with M_PI/180 is formed in double and
// narrowed, as the float *= double of the reference does.
float arc_rad(float anarc, int points) {
  float r = anarc;
  if (r < 1.0f) r *= static_cast<float>(points - 1);
  return static_cast<float>(static_cast<double>(r) * (M_PI / 180));
}

// angO (lct.cpp:478-484): an arc above one degree is divided into steps.
float step_rad(float anarc, int points) {
  float r = anarc;
  if (r > 1.0f && points > 1) r /= static_cast<float>(points - 1);
  return static_cast<float>(static_cast<double>(r) * (M_PI / 180));
}

// lct.cpp:512-517: theta_i = i*arc/N in float; entry {sin, cos} (.x = sin).
void angle_table(int n, float anarc, float *sincos) {
  const float arc = arc_rad(anarc, n);
  for (int i = 0; i < n; ++i) {
    const float th = (static_cast<float>(i) * arc) / static_cast<float>(n);
    sincos[2 * i] = std::sin(th);
    sincos[2 * i + 1] = std::cos(th);
  }
}

// FP64 angle table: theta_i = i*arc/N with the arc in double
// (anarc below one degree is a step, as in arcO).
void angle_table64(int n, float anarc, double *sincos) {
  double a = anarc;
  if (a < 1.0) a *= (n - 1);
  a *= M_PI / 180.0;
  for (int i = 0; i < n; ++i) {
    const double th = i * a / n;
    sincos[2 * i] = std::sin(th);
    sincos[2 * i + 1] = std::cos(th);
  }
}

// Angles whose FP64 entries have equal sines and opposite cosines (to
// 1e-12) pair up; an angle with cosine 0 is its own mirror and stays
// single. Lookup by the entries rounded to 1e-9, each candidate checked
// exactly; a pair missed at a rounding boundary only costs speed.
int mirror_slots(const std::vector<double> &sc64, std::vector<MirrorSlot> &slots) {
  const int n = static_cast<int>(sc64.size() / 2);
  auto key = [](double s, double c) { return std::make_pair(std::llround(s * 1e9), std::llround(c * 1e9)); };
  std::map<std::pair<long long, long long>, std::vector<int>> by_key;
  for (int i = 0; i < n; ++i) by_key[key(sc64[2 * i], sc64[2 * i + 1])].push_back(i);
  std::vector<int> partner(n, -1);
  for (int i = 0; i < n; ++i) {
    if (partner[i] >= 0) continue;
    const double s = sc64[2 * i], c = sc64[2 * i + 1];
    const auto f = by_key.find(key(s, -c));
    if (f == by_key.end()) continue;
    for (int j : f->second) {
      if (j == i || partner[j] >= 0) continue;
      if (std::fabs(sc64[2 * j] - s) <= 1e-12 && std::fabs(sc64[2 * j + 1] + c) <= 1e-12) {
        partner[i] = j;
        partner[j] = i;
        break;
      }
    }
  }
  slots.clear();
  for (int i = 0; i < n; ++i)
    if (partner[i] > i) slots.push_back(MirrorSlot{sc64[2 * i], sc64[2 * i + 1], i, partner[i], 0, 0});
  const int paired = static_cast<int>(slots.size());
  for (int i = 0; i < n; ++i)
    if (partner[i] < 0) slots.push_back(MirrorSlot{sc64[2 * i], sc64[2 * i + 1], i, -1, 0, 0});
  return paired;
}

namespace {

// In-place iterative complex FFT (forward, unnormalised), n a power of two.
void fft_inplace(std::vector<std::complex<double>> &a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const size_t half = len / 2;
    for (size_t k = 0; k < half; ++k) {
      const double ang = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(len);
      const std::complex<double> wk(std::cos(ang), std::sin(ang));
      for (size_t i = k; i < n; i += len) {
        const std::complex<double> x = a[i + half] * wk;
        a[i + half] = a[i] - x;
        a[i] += x;
      }
    }
  }
}

double bessel_i0(double x) {
  // power series sum_k (x^2/4)^k / (k!)^2
  double sum = 1.0, term = 1.0;
  const double q = 0.25 * x * x;
  for (int k = 1; k < 1000; ++k) {
    term *= q / (static_cast<double>(k) * k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

}  // namespace

// Filter::fill (lct.cpp:122-209), then the CTrec gain zidth/w and the REF
// zero-frequency rule (lct.cpp:505-508). The spatial Ram-Lak taps
// h[0] = 1/4, h[odd n < Z/2] = -1/(n pi)^2 are rounded to float as in the
// reference; their half-complex inverse transform H[k] = h0 + 2 sum h[n]
// cos(2 pi n k / Z) is evaluated in FP64 (as the DFT of the even extension)
// and rounded once. The window factor of each filter follows the float /
// double mix of lct.cpp:146-205, including its PARZEN, LANCKZOS and
// NONE == RAMP behaviour.
// Filter::fill (lct.cpp:122-209) for a line of n points: the Ram-Lak kernel
// (h[0] = 1/4, h[odd k < n/2] = -1/(k pi)^2) taken to the frequency domain
// by a half-complex inverse transform (H[k] = h0 + 2 sum h[m] cos(2 pi m k/n);
// FP64 FFT for powers of two, a direct FP64 sum otherwise), times the
// window of f = 2 min(k, n - k)/n, halved.
int filter_fill(int filter, float param, int n, std::vector<float> &W) {
  if (n < 1) return fail(CTAS_E_ARG, "Zero-size filter requested.");
  if (filter < CTAS_FILTER_NONE || filter > CTAS_FILTER_GAUSS)
    return fail(CTAS_E_ARG, "filter window: unknown filter code " + std::to_string(filter));
  if (filter == CTAS_FILTER_KAISER && param < 0.0f)
    return fail(CTAS_E_ARG, "Alpha parameter of Kaiser filter must be non-negative (set " +
                                std::to_string(param) + ").");
  if (filter == CTAS_FILTER_GAUSS && param <= 0.0f)
    return fail(CTAS_E_ARG, "Sigma parameter of Guass filter must be strictly positive (set " +
                                std::to_string(param) + ").");
  const int Z = n;
  std::vector<std::complex<double>> a(Z, 0.0);
  a[0] = static_cast<double>(0.25f);
  for (int m = 1; m < Z / 2; m += 2) {
    const float h = static_cast<float>(-1.0 / (static_cast<double>(m) * m * M_PI * M_PI));
    a[m] = h;
    a[Z - m] = h;
  }
  if ((Z & (Z - 1)) == 0) {
    fft_inplace(a);
  } else {
    std::vector<std::complex<double>> b(Z);
    for (int k = 0; k < Z; ++k) {
      double acc = a[0].real();
      for (int m = 1; m < Z / 2; m += 2)
        acc += 2.0 * a[m].real() * std::cos(2.0 * M_PI * static_cast<double>((static_cast<long long>(m) * k) % Z) / Z);
      b[k] = acc;
    }
    a.swap(b);
  }
  W.assign(Z, 0.0f);
  for (int k = 0; k < Z; ++k) {
    const float H = static_cast<float>(a[k].real());
    float fp = static_cast<float>(k < Z / 2 ? k : Z - k);
    fp = static_cast<float>(static_cast<double>(fp) * (2.0 / Z));
    switch (filter) {
      case CTAS_FILTER_NONE:
      case CTAS_FILTER_RAMP: fp = 1.0f; break;
      case CTAS_FILTER_BARLETT: fp = 1 - fp; break;
      case CTAS_FILTER_WELCH: fp = 1 - fp * fp; break;
      case CTAS_FILTER_PARZEN:
        fp = (fp <= 0.5f) ? 1 - 6 * fp * fp + 3 * fp * fp * fp : 2 * (1 + fp * fp * fp);
        break;
      case CTAS_FILTER_HANN: fp = static_cast<float>(0.5 * (1 + std::cos(M_PI * fp))); break;
      case CTAS_FILTER_HAMMING: fp = static_cast<float>(0.54 + 0.46 * std::cos(M_PI * fp)); break;
      case CTAS_FILTER_BLACKMAN:
        fp = static_cast<float>(0.42 + 0.5 * std::cos(M_PI * fp) + 0.08 * std::cos(2 * M_PI * fp));
        break;
      case CTAS_FILTER_LANCKZOS:
        fp = (fp == 0.0f) ? 0.0f : static_cast<float>(std::fabs(std::sin(M_PI * fp)) / fp);
        break;
      case CTAS_FILTER_KAISER: {
        const float arg = param * std::sqrt(1 - fp * fp);
        fp = static_cast<float>(bessel_i0(arg) / bessel_i0(param));
        break;
      }
      case CTAS_FILTER_GAUSS: fp = std::pow(2.0f, (-fp) * fp / (param * param)); break;
    }
    W[k] = static_cast<float>(static_cast<double>(H) * (static_cast<double>(fp) / 2.0));
  }
  return CTAS_OK;
}

// CTrec's window (lct.cpp:487-508): Filter::fill at the padded width Z,
// times the gain Z/w; REF: no DC.
int filter_window(int filter, float param, int w, int contrast, std::vector<float> &W) {
  if (w < 1) return fail(CTAS_E_ARG, "filter window: width must be positive");
  const int Z = zidth(w);
  if (int rc = filter_fill(filter, param, Z, W)) return rc;
  const float gain = static_cast<float>(Z) / static_cast<float>(w);
  for (float &v : W) v = v * gain;
  if (contrast == CTAS_CONTRAST_REF) W[0] = 0.0f;
  return CTAS_OK;
}

// ts_add's filter (lct.cpp:758-777) convolves each row of n samples
// circularly with g = IDFT_n(W), W = Filter::fill(n), by FFTW transforms of
// exactly n points. The device transforms powers of two only, so for any
// other n the same circular convolution runs as a zero-padded one of
// Z = zidth(n) >= 2n - 1 points: the 2n - 1 lags -(n-1)..n-1 of the
// periodic g then fall on distinct residues mod Z, and the window of that
// longer line is DFT_Z of g laid out at lags 0..n-1 and Z-(n-1)..Z-1 (real,
// as g is even). For a power of two n >= 2, Z = n and the window is W itself.
// Both windows are for a transform whose inverse is scaled by 1/Z.
int ts_window(int filter, float param, int n, int &Z, std::vector<float> &out) {
  std::vector<float> W;
  if (int rc = filter_fill(filter, param, n, W)) return rc;
  if (n >= 2 && (n & (n - 1)) == 0) {
    Z = n;
    out = W;
    return CTAS_OK;
  }
  Z = std::max(2, zidth(n));
  // g[m] = (1/n) (sum_k hc[k] cos(2 pi k m / n) terms of FFTW's HC2R with a
  // zero imaginary half): W is even, so g[m] = (1/n) sum_k W[k] cos(2 pi k m/n).
  std::vector<double> cs(n);
  for (int j = 0; j < n; ++j) cs[j] = std::cos(2.0 * M_PI * j / n);
  std::vector<std::complex<double>> a(Z, 0.0);
  for (int m = 0; m < n; ++m) {
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc += static_cast<double>(W[k]) * cs[(static_cast<long long>(k) * m) % n];
    acc /= n;
    a[m] += acc;
    if (m) a[Z - m] += acc;
  }
  fft_inplace(a);
  out.resize(Z);
  for (int k = 0; k < Z; ++k) out[k] = static_cast<float>(a[k].real());
  return CTAS_OK;
}

bool ts_odd_basis(int filter, float param, int n, int Z, std::vector<float2> &uv, float &gain) {
  uv.clear();
  gain = 0.0f;
  if (n < 3 || (n & 1) == 0) return false;
  std::vector<float> W;
  if (filter_fill(filter, param, n, W) != CTAS_OK) return false;
  const int m = (n - 1) / 2;
  gain = static_cast<float>(static_cast<double>(Z) * (static_cast<double>(W[m]) - W[m + 1]) / n);
  uv.resize(n);
  for (int x = 0; x < n; ++x) {
    const double sg = (x & 1) ? -1.0 : 1.0, a = M_PI * x / n;
    uv[x] = make_float2(static_cast<float>(sg * std::cos(a)), static_cast<float>(sg * std::sin(a)));
  }
  return true;
}

}  // namespace ctas
