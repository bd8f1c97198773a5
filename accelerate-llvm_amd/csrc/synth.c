/* Seeded synthetic inputs of the benchmark configurations (host only).
 *
 * One stream per array: xoshiro256** (Blackman and Vigna, 2018) whose four
 * state words are the first four outputs of splitmix64 started at the seed.
 * Every generator is a pure function of its arguments, so the benchmark, the
 * tests and the CPU baseline see the same arrays.  Uniform doubles take the
 * top 53 bits of an output (scaled by 2^-53, [0, 1)); normals come from the
 * Box-Muller transform of two consecutive uniforms (cosine then sine).
 *
 * Built into its own library (lib/libacc_synth.so): it is input synthesis
 * for the benchmark and the tests, not part of the skeleton interface.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  uint64_t s[4];
} acc_xo;

static uint64_t sm64_next(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static acc_xo xo_from_seed(uint64_t seed) {
  acc_xo r;
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r.s[i] = sm64_next(&x);
  return r;
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

static inline uint64_t xo_next(acc_xo* r) {
  uint64_t* s = r->s;
  const uint64_t out = rotl64(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return out;
}

static inline double xo_unit(acc_xo* r) { return (double)(xo_next(r) >> 11) * 0x1.0p-53; }

/* n splitmix64 outputs from state x (known-answer checks) */
void acc_synth_splitmix64(uint64_t x, uint64_t* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = sm64_next(&x);
}

/* n xoshiro256** outputs from the given state (known-answer checks) */
void acc_synth_xoshiro(const uint64_t* state, uint64_t* out, int64_t n) {
  acc_xo r = {{state[0], state[1], state[2], state[3]}};
  for (int64_t i = 0; i < n; ++i) out[i] = xo_next(&r);
}

/* Int32 uniform on the closed range [lo, hi] (output mod the span) */
void acc_synth_uniform_i32(uint64_t seed, int32_t lo, int32_t hi, int32_t* out, int64_t n) {
  acc_xo r = xo_from_seed(seed);
  const uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
  for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)((int64_t)lo + (int64_t)(xo_next(&r) % span));
}

/* Float32 uniform on [lo, hi) from the top 24 bits of each output */
void acc_synth_uniform_f32(uint64_t seed, float lo, float hi, float* out, int64_t n) {
  acc_xo r = xo_from_seed(seed);
  for (int64_t i = 0; i < n; ++i) {
    const float u = (float)((double)(xo_next(&r) >> 40) * 0x1.0p-24);
    out[i] = lo + (hi - lo) * u;
  }
}

/* Float32 N(0, 1): Box-Muller over pairs (u1 in (0, 1], u2 in [0, 1)) */
void acc_synth_normal_f32(uint64_t seed, float* out, int64_t n) {
  acc_xo r = xo_from_seed(seed);
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    const double u1 = 1.0 - xo_unit(&r), u2 = xo_unit(&r);
    const double rad = sqrt(-2.0 * log(u1));
    out[i] = (float)(rad * cos(two_pi * u2));
    if (i + 1 < n) out[i + 1] = (float)(rad * sin(two_pi * u2));
  }
}

/* bin indices in [0, nbins) with P(k) proportional to (k+1)^-s: the first
 * bin whose cumulative probability exceeds a uniform double */
int acc_synth_zipf_i32(uint64_t seed, double s, int32_t nbins, int32_t* out, int64_t n) {
  if (nbins <= 0) return -1;
  double* cdf = (double*)malloc(sizeof(double) * (size_t)nbins);
  if (!cdf) return -1;
  double total = 0.0;
  for (int32_t k = 0; k < nbins; ++k) {
    total += pow((double)(k + 1), -s);
    cdf[k] = total;
  }
  for (int32_t k = 0; k < nbins; ++k) cdf[k] /= total;
  acc_xo r = xo_from_seed(seed);
  for (int64_t i = 0; i < n; ++i) {
    const double u = xo_unit(&r);
    int32_t a = 0, b = nbins - 1;
    while (a < b) {
      const int32_t m = a + (b - a) / 2;
      if (cdf[m] > u) b = m;
      else a = m + 1;
    }
    out[i] = a;
  }
  free(cdf);
  return 0;
}

/* nseg Int64 lengths: L_i = exp(u_i) with u_i ~ U(0, ln maxlen), scaled to
 * sum to total and rounded (each at least 1).  A shortfall goes to the first
 * longest segment; an excess (from the floor of 1) is taken off the segments
 * in index order, none below 1.  Needs 1 <= nseg <= total. */
int acc_synth_segment_lengths(uint64_t seed, int64_t nseg, int64_t maxlen, int64_t total, int64_t* out) {
  if (nseg <= 0 || nseg > total || maxlen < 1) return -1;
  double* len = (double*)malloc(sizeof(double) * (size_t)nseg);
  if (!len) return -1;
  acc_xo r = xo_from_seed(seed);
  const double lnmax = log((double)maxlen);
  double sum = 0.0;
  for (int64_t i = 0; i < nseg; ++i) {
    len[i] = exp(xo_unit(&r) * lnmax);
    sum += len[i];
  }
  int64_t got = 0, longest = 0;
  for (int64_t i = 0; i < nseg; ++i) {
    int64_t v = (int64_t)llround(len[i] * (double)total / sum);
    out[i] = v < 1 ? 1 : v;
    got += out[i];
    if (out[i] > out[longest]) longest = i;
  }
  if (got <= total) {
    out[longest] += total - got;
  } else {
    int64_t excess = got - total;
    for (int64_t i = 0; i < nseg && excess > 0; ++i) {
      const int64_t give = out[i] - 1 < excess ? out[i] - 1 : excess;
      out[i] -= give;
      excess -= give;
    }
  }
  free(len);
  return 0;
}
