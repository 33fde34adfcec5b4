/*
 * twiddle.c -- DFT roots of unity for the device twiddle tables.
 *
 * exp(sign * 2*pi*i * m / N) is computed in x87 long double (64-bit
 * mantissa) after an exact integer reduction to the first quadrant:
 *   4m = q*N + r,  0 <= r < N,  angle = q*pi/2 + (pi/2)*(r/N),
 * and for r > N/2 the complementary angle (pi/2)*((N-r)/N) is used with
 * sine and cosine swapped.  The result is rounded once to double, so the
 * values at multiples of pi/4 and the symmetries W^(N-m) = conj(W^m) hold
 * exactly.
 */
#include <math.h>
#include "host.h"

static void quarter_sincos(long r, long N, long double *c, long double *s) {
  /* (pi/2)*(r/N) with 0 <= r < N */
  static const long double half_pi = 1.57079632679489661923132169163975144L;
  if (2 * r > N) {
    const long double a = half_pi * (long double)(N - r) / (long double)N;
    *c = sinl(a);
    *s = cosl(a);
  } else if (2 * r == N) {
    /* pi/4: both sqrt(1/2) */
    *c = *s = sqrtl(0.5L);
  } else {
    const long double a = half_pi * (long double)r / (long double)N;
    *c = cosl(a);
    *s = sinl(a);
  }
}

void pa_twiddles(long N, long count, int sign, double *out) {
  pa_twiddles_step(N, count, 1, sign, out);
}

/* exp(sign*2*pi*i*step*j/N) for j in [0, count) */
void pa_twiddles_step(long N, long count, long step, int sign, double *out) {
  for (long j = 0; j < count; j++) {
    long mm = (long)(((long long)step * j) % N);
    if (mm < 0) mm += N;
    const long long four_m = 4LL * mm;
    const long q = (long)(four_m / N);
    const long r = (long)(four_m % N);
    long double c, s, cc, ss;
    quarter_sincos(r, N, &c, &s);
    switch (q & 3) {
      case 0: cc = c; ss = s; break;
      case 1: cc = -s; ss = c; break;
      case 2: cc = -c; ss = -s; break;
      default: cc = s; ss = -c; break;
    }
    out[2 * j] = (double)cc;
    out[2 * j + 1] = (double)(sign < 0 ? -ss : ss);
  }
}

int64_t pfk_max_grid(int threads) {
  if (threads < 1) threads = 1;
  return (int64_t)(0xFFFFFFFFu / (unsigned)threads);
}
