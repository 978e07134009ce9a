/* Exhaustive and randomized checks of the exact arithmetic rewrites used by
 * the pull-push kernel, against plain IEEE float operations (the form the
 * FP32 oracle uses). Built and run by tests/test_arith_identities.py. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* x / b from y = RN(1/b): q = RN(x y), r = fma(-q, b, x), RN(q + r y). */
static float div_by_recip(float x, float b, float y) {
  volatile float q = x * y;
  const float r = fmaf(-q, b, x);
  return fmaf(r, y, q);
}

/* Mismatches of div_by_recip against x / b for every significand of x in the
 * binades [2^e0, 2^e1), both signs; y = RN(1/b) moved by y_ulps ulps (0 for
 * the kernel's form, nonzero to show that the harness catches a bad y). */
long check_div_by_recip(float b, int y_ulps, int e0, int e1) {
  volatile float bb = b;
  const float y = u2f(f2u(1.0f / bb) + (uint32_t)y_ulps);
  long bad = 0;
  for (int e = e0; e < e1; ++e)
    for (uint32_t m = 0; m < (1u << 23); ++m)
      for (uint32_t s = 0; s < 2; ++s) {
        const float x = u2f((s << 31) | ((uint32_t)(e + 127) << 23) | m);
        volatile float ref = x / bb;
        if (f2u(div_by_recip(x, b, y)) != f2u(ref)) ++bad;
      }
  return bad;
}

/* Bytes c whose unpackUnorm4x8 value c / 255 differs from the splat's form
 * div_by_recip(c, 255, RN(1/255)); with product_only, from RN(c RN(1/255))
 * instead (which the harness must catch). */
long check_unorm8(int product_only) {
  volatile float d = 255.0f;
  const float y = 1.0f / d;
  long bad = 0;
  for (uint32_t c = 0; c < 256; ++c) {
    const float x = (float)c;
    volatile float ref = x / d;
    const float got = product_only ? x * y : div_by_recip(x, 255.0f, y);
    if (f2u(got) != f2u(ref)) ++bad;
  }
  return bad;
}

/* m^2 with m the midpoint between R and the next float (exact in double);
 * midpoint = 0 gives the naive bound R^2 instead. */
static double sqrt_gt_bound(float R, int midpoint) {
  const float n = u2f(f2u(R) + 1u);
  const double m = midpoint ? ((double)R + (double)n) * 0.5 : (double)R;
  return m * m;
}

static uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Mismatches of (double)x > sqrt_gt_bound(R) against sqrtf(x) > R for n
 * random radii R in [2^-8, 2^5) and, for each, the 65 floats x around
 * RN(R * R) (where the decisions change) plus one random x. */
long check_sqrt_bound(long n, uint64_t seed, int midpoint) {
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    const uint32_t e = 119u + (uint32_t)(splitmix(&seed) % 13u);
    const float R = u2f((e << 23) | (uint32_t)(splitmix(&seed) & 0x7FFFFFu));
    const double b = sqrt_gt_bound(R, midpoint);
    const uint32_t c = f2u(R * R);
    for (int k = -33; k <= 32; ++k) {
      const float x = k < 32 ? u2f(c + (uint32_t)k) : R * R * (float)(splitmix(&seed) % 1000u) / 500.0f;
      volatile float sx = sqrtf(x);
      const int ref = sx > R;
      const int got = (double)x > b;
      if (ref != got) ++bad;
    }
  }
  return bad;
}

/* RD(m^2) in FP32 alone: m^2 = R * next(R) + u^2 / 4 (u = ulp(R)), and the
 * added u^2 / 4 is below one ulp of the product, so RD(m^2) = RD(R next(R)):
 * the product rounded to nearest, moved down one float when its FMA error
 * shows it was rounded up. */
static float rd_bound(float R) {
  const float n = u2f(f2u(R) + 1u);
  volatile float p = R * n;
  const float e = fmaf(R, n, -p);
  return e < 0.0f ? u2f(f2u(p) - 1u) : p;
}

/* RD(m^2) from the exact FP64 m^2. */
static float rd_bound_f64(float R) {
  const double b = sqrt_gt_bound(R, 1);
  float f = (float)b;
  if ((double)f > b) f = u2f(f2u(f) - 1u);
  return f;
}

/* Mismatches of rd_bound against rd_bound_f64 for every float R in the
 * binades [2^e0, 2^e1), plus 0 and 4096 denormals. */
long check_rd_bound(int e0, int e1) {
  long bad = 0;
  for (int e = e0; e < e1; ++e)
    for (uint32_t m = 0; m < (1u << 23); ++m) {
      const float R = u2f(((uint32_t)(e + 127) << 23) | m);
      if (f2u(rd_bound(R)) != f2u(rd_bound_f64(R))) ++bad;
    }
  for (uint32_t d = 0; d < 4096; ++d) {
    const float R = u2f(d * 2047u);
    if (f2u(rd_bound(R)) != f2u(rd_bound_f64(R))) ++bad;
  }
  return bad;
}

/* Mismatches of x > rd_bound(R) against sqrtf(x) > R, as check_sqrt_bound. */
long check_rd_bound_sqrt(long n, uint64_t seed) {
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    const uint32_t e = 119u + (uint32_t)(splitmix(&seed) % 13u);
    const float R = u2f((e << 23) | (uint32_t)(splitmix(&seed) & 0x7FFFFFu));
    const float b = rd_bound(R);
    const uint32_t c = f2u(R * R);
    for (int k = -33; k <= 32; ++k) {
      const float x = k < 32 ? u2f(c + (uint32_t)k) : R * R * (float)(splitmix(&seed) % 1000u) / 500.0f;
      volatile float sx = sqrtf(x);
      if ((sx > R) != (x > b)) ++bad;
    }
  }
  return bad;
}

/* Pairs (s, t) where the splat's sign test s < -t differs from RN(s + t) < 0
 * (the dot product's last addition): random floats of every magnitude, both
 * signs, with exact cancellations (t = -s and neighbours of -s), signed
 * zeros and subnormals mixed in. With naive = 1 the test is s < t instead,
 * which the harness must catch. */
long check_sign_of_sum(long n, uint64_t seed, int naive) {
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    const uint64_t r = splitmix(&seed);
    float s = u2f((uint32_t)r), t;
    switch ((r >> 32) & 7u) {
      case 0: t = -s; break;                                    /* exact 0 */
      case 1: t = u2f(f2u(-s) + 1u); break;                     /* one ulp off */
      case 2: t = u2f(f2u(-s) - 1u); break;
      case 3: t = u2f((uint32_t)(r >> 32) & 0x807FFFFFu); break; /* subnormal */
      case 4: t = ((r >> 35) & 1u) ? 0.0f : -0.0f; break;
      default: t = u2f((uint32_t)(splitmix(&seed))); break;
    }
    if (isnan(s) || isnan(t) || isinf(s) || isinf(t)) continue;
    volatile float sum = s + t;
    const int ref = sum < 0.0f;
    const int got = naive ? (s < t) : (s < -t);
    if (ref != got) ++bad;
  }
  return bad;
}
