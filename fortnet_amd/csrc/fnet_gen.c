/*
 * fnet_gen.c -- host helper of the synthetic workload generator (fortnet_amd/workloads.py):
 * random sequential addition of points in a cubic periodic box with a minimum distance, the
 * generator of the random Si and ternary cells.
 *
 * Bit-identical to the numpy restatement in workloads._rsa: candidates come from the same
 * counter-based SplitMix64 stream (3 x 64 uniforms per batch of 64 candidates, uniform =
 * (z >> 11) 2^-53), the minimum-image difference is d - L rint(d / L) with ties to even (as
 * numpy.round), squared distances are summed as (dx^2 + dy^2) + dz^2, and products are not
 * contracted into FMAs.  A candidate is accepted when every accepted point lies at least dmin
 * away (the first candidate always).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static const uint64_t kGamma = 0x9E3779B97F4A7C15ull;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double uniform_at(uint64_t seed, uint64_t idx) {
  const uint64_t z = mix64(seed + idx * kGamma);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static double min_image(double d, double L) { return d - L * rint(d / L); }

static int too_close(const double* pts, int a, double x, double y, double z, double L,
                     double d2min) {
  const double dx = min_image(pts[3 * a] - x, L);
  const double dy = min_image(pts[3 * a + 1] - y, L);
  const double dz = min_image(pts[3 * a + 2] - z, L);
  double d2 = dx * dx;
  d2 = d2 + dy * dy;
  d2 = d2 + dz * dz;
  return d2 < d2min;
}

static int bin_of(double x, double L, int nb) {
  int b = (int)(x / L * nb);
  return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

/* Fills pts[n][3]; the stream is at counter *k (uniforms drawn so far) and is advanced.
 * Accepted points are kept in a periodic grid of nb^3 bins at least dmin wide, so only the 27
 * bins around a candidate can hold a point closer than dmin (all points are checked when fewer
 * than 3 bins fit).  Returns 0, or -1 when max_tries candidates were not enough (or no memory). */
int fnet_rsa(uint64_t seed, uint64_t* k, int n, double L, double dmin, long max_tries,
             double* pts) {
  const double d2min = dmin * dmin;
  int nb = dmin > 0.0 ? (int)floor(L / dmin) : 1;
  if (nb > 64) nb = 64;
  const int grid = nb >= 3;
  int* head = NULL;
  int* next = NULL;
  if (grid) {
    head = (int*)malloc(sizeof(int) * (size_t)nb * nb * nb);
    next = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!head || !next) {
      free(head);
      free(next);
      return -1;
    }
    for (int c = 0; c < nb * nb * nb; ++c) head[c] = -1;
  }
  int acc = 0;
  long tries = 0;
  int rc = 0;
  while (acc < n) {
    const uint64_t base = *k;
    *k += 3 * 64;
    for (int c = 0; c < 64 && acc < n; ++c) {
      if (++tries > max_tries) {
        rc = -1;
        goto done;
      }
      const double x = uniform_at(seed, base + 1 + 3 * c) * L;
      const double y = uniform_at(seed, base + 2 + 3 * c) * L;
      const double z = uniform_at(seed, base + 3 + 3 * c) * L;
      int ok = 1;
      if (grid) {
        const int bx = bin_of(x, L, nb), by = bin_of(y, L, nb), bz = bin_of(z, L, nb);
        for (int d = 0; d < 27 && ok; ++d) {
          const int cx = (bx + d / 9 - 1 + nb) % nb, cy = (by + (d / 3) % 3 - 1 + nb) % nb,
                    cz = (bz + d % 3 - 1 + nb) % nb;
          for (int a = head[(cx * nb + cy) * nb + cz]; a >= 0 && ok; a = next[a])
            if (too_close(pts, a, x, y, z, L, d2min)) ok = 0;
        }
      } else {
        for (int a = 0; a < acc && ok; ++a)
          if (too_close(pts, a, x, y, z, L, d2min)) ok = 0;
      }
      if (ok) {
        pts[3 * acc] = x;
        pts[3 * acc + 1] = y;
        pts[3 * acc + 2] = z;
        if (grid) {
          const int cell = (bin_of(x, L, nb) * nb + bin_of(y, L, nb)) * nb + bin_of(z, L, nb);
          next[acc] = head[cell];
          head[cell] = acc;
        }
        ++acc;
      }
    }
  }
done:
  free(head);
  free(next);
  return rc;
}
