/* Counter-based random streams of the MI355X Smoldyn engine.
 *
 * The reference draws every random number from one global SFMT-19937 stream
 * in program order (source/libSteve/random2.h:21-58), which serialises the
 * whole time step.  The device engine instead derives each draw from
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11) keyed by the simulation seed and a counter built from
 * (molecule serial number, step index, purpose).  Any thread can then compute
 * the draws of any molecule without coordination, and the CPU checker computes
 * the very same values, so whole trajectories can be compared bit for bit.
 *
 * This header is plain C, usable from host C/C++ and from HIP device code.
 */
#ifndef SMOLDYN_AMD_RNG_H
#define SMOLDYN_AMD_RNG_H

#include <stdint.h>
#if !defined(__HIPCC__) && !defined(__HIP__)
#include <math.h>
#endif

#if defined(__HIPCC__) || defined(__HIP__)
#define SMOLRNG_FN __host__ __device__ static inline
#else
#define SMOLRNG_FN static inline
#endif

/* Counter word 3 carries the purpose of a draw in its low 8 bits and a
   sub-index (reaction number, attempt number, wall number) above them. */
enum SmolRngPurpose {
  SMOLRNG_DIFFUSE = 1,   /* Gaussian table indices for one diffusion step */
  SMOLRNG_UNIRXN = 2,    /* first-order reaction coin flips, 4 reactions per block */
  SMOLRNG_PRODDIR = 3,   /* product unbinding direction, 2 attempts per block */
  SMOLRNG_BIRXN = 4,     /* bimolecular reaction probability < 1 */
  SMOLRNG_WALL = 5,      /* absorbing-wall coin flip */
  SMOLRNG_SURFACE = 6,   /* surface-action draws of one molecule, sub = pass<<8 | iteration */
  SMOLRNG_PLACE = 7,     /* initial uniform placement */
  SMOLRNG_RXNSURF = 8,   /* surface-action draws of a reacting pair (rxnXsurface) */
  SMOLRNG_ZERO = 9,      /* zeroth-order reaction count: c0 = reaction, c1 = chunk, sub = block */
  SMOLRNG_ZEROPOS = 10,  /* zeroth-order product position: c0 = reaction, c1 = event */
  SMOLRNG_PRODROT = 11   /* random rotation of product offset vectors, sub = attempt */
};

/* Surface draw passes (bits 8-15 of the SMOLRNG_SURFACE sub-index). */
enum SmolRngSurfPass {
  SMOLRNG_SURF_DIFFUSED = 0,  /* collisions after diffusion, sub = iteration 1..100 */
  SMOLRNG_SURF_REBORN = 1,    /* collisions of reaction products */
  SMOLRNG_SURF_DESORBED = 2,  /* collisions of a desorbed molecule */
  SMOLRNG_SURF_BOUND = 3,     /* surface-bound action (desorption) */
  SMOLRNG_SURF_DESORBDIST = 4 /* desorption distance */
};

typedef struct { uint32_t v[4]; } smolrng4;

SMOLRNG_FN uint32_t smolrng_mulhilo(uint32_t a, uint32_t b, uint32_t* hi) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  return (uint32_t)p;
}

/* Philox4x32 with 10 rounds. */
SMOLRNG_FN smolrng4 smolrng_philox(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1,
                                   uint32_t c2, uint32_t c3) {
  smolrng4 r;
  int i;
  for (i = 0; i < 10; i++) {
    uint32_t hi0, hi1;
    uint32_t lo0 = smolrng_mulhilo(0xD2511F53u, c0, &hi0);
    uint32_t lo1 = smolrng_mulhilo(0xCD9E8D57u, c2, &hi1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  r.v[0] = c0;
  r.v[1] = c1;
  r.v[2] = c2;
  r.v[3] = c3;
  return r;
}

/* Number of events of a zeroth-order reaction with mean lambda per step is
   the sum over chunks j < m = ceil(lambda / SMOLRNG_ZERO_CHUNK) of Poisson
   draws of mean mu = lambda / m.  One chunk: multiply uniforms until the
   product falls to g = exp(-mu) or below (the multiplication method of the
   reference's poisrandD for small means); g is computed once on the host. */
#define SMOLRNG_ZERO_CHUNK 8.0
SMOLRNG_FN int smolrng_poisson_chunk(uint32_t k0, uint32_t k1, uint32_t rxn, uint32_t chunk, uint32_t step,
                                     double g) {
  uint32_t blk = 0;
  int idx = 1, n = 0;
  smolrng4 b = smolrng_philox(k0, k1, rxn, chunk, step, SMOLRNG_ZERO);
  double t = (double)b.v[0] * (1.0 / 4294967296.0);
  while (t > g) {
    if (idx == 4) {
      blk++;
      b = smolrng_philox(k0, k1, rxn, chunk, step, SMOLRNG_ZERO | (blk << 8));
      idx = 0;
    }
    t *= (double)b.v[idx++] * (1.0 / 4294967296.0);
    n++;
  }
  return n;
}

/* Draw block for one molecule (serial number) at one step. */
SMOLRNG_FN smolrng4 smolrng_mol(uint32_t k0, uint32_t k1, uint64_t serial, uint32_t step,
                                uint32_t purpose, uint32_t sub) {
  return smolrng_philox(k0, k1, (uint32_t)serial, (uint32_t)(serial >> 32), step,
                        (purpose & 0xFFu) | (sub << 8));
}

/* Same as the reference's randCOD (genrand_real2): u32 * 2^-32, in [0,1). */
SMOLRNG_FN double smolrng_u01(uint32_t v) { return (double)v * (1.0 / 4294967296.0); }

/* Same as the reference's randCCD (genrand_real1): u32 / (2^32 - 1), in [0,1]. */
SMOLRNG_FN double smolrng_u01cc(uint32_t v) { return (double)v * (1.0 / 4294967295.0); }
/* Draw block for a pair of molecules at one step. */
SMOLRNG_FN smolrng4 smolrng_pair(uint32_t k0, uint32_t k1, uint64_t ser1, uint64_t ser2, uint32_t step,
                                 uint32_t purpose, uint32_t sub) {
  return smolrng_philox(k0, k1, (uint32_t)ser1, (uint32_t)ser2, step, (purpose & 0xFFu) | (sub << 8));
}
/* Same as the reference's randOOD (genrand_real3): (u32 + 0.5) * 2^-32, in (0,1). */
SMOLRNG_FN double smolrng_u01open(uint32_t v) { return ((double)v + 0.5) * (1.0 / 4294967296.0); }

/* Uniform in [-1,1) from 32 bits, exact in double. */
SMOLRNG_FN double smolrng_sym(uint32_t v) { return (double)v * (1.0 / 2147483648.0) - 1.0; }

/* Uniformly random unit vector by Marsaglia's rejection method (Ann. Math.
   Stat. 43:645, 1972), using only +,*,/ and sqrt so that host and device
   results agree bit for bit.  Each Philox block yields two attempts; attempt
   blocks are numbered by sub.  dim is 1, 2 or 3. */
SMOLRNG_FN void smolrng_unitvec(uint32_t k0, uint32_t k1, uint64_t serial, uint32_t step,
                                int dim, double* u) {
  uint32_t a;
  if (dim == 1) {
    smolrng4 r = smolrng_mol(k0, k1, serial, step, SMOLRNG_PRODDIR, 0);
    u[0] = (r.v[0] & 1u) ? 1.0 : -1.0;
    return;
  }
  for (a = 0; a < 64; a++) {
    smolrng4 r = smolrng_mol(k0, k1, serial, step, SMOLRNG_PRODDIR, a);
    int t;
    for (t = 0; t < 2; t++) {
      double x1 = smolrng_sym(r.v[2 * t]);
      double x2 = smolrng_sym(r.v[2 * t + 1]);
      double s = x1 * x1 + x2 * x2;
      if (s > 0.0 && s < 1.0) {
        if (dim == 2) {
          u[0] = (x1 * x1 - x2 * x2) / s;
          u[1] = (2.0 * x1 * x2) / s;
        } else {
          double q = 2.0 * sqrt(1.0 - s);
          u[0] = x1 * q;
          u[1] = x2 * q;
          u[2] = 1.0 - 2.0 * s;
        }
        return;
      }
    }
  }
  /* 128 rejected attempts has probability (1-pi/4)^128 < 1e-85. */
  u[0] = 1.0;
  u[1] = 0.0;
  if (dim == 3) u[2] = 0.0;
}

/* Uniformly distributed rotation (Haar measure) for product offset vectors
   off the x axis: a unit quaternion by Marsaglia's method (two points drawn
   in the unit disk, Ann. Math. Stat. 43:645, 1972) turned into a rotation
   matrix m (row major).  Arithmetic and sqrt only, so host and device agree
   to the bit.  The reference draws Euler angles of the same distribution
   (DirCosMD with thetarandCCD, smolreact.c:2846-2848). */
SMOLRNG_FN void smolrng_rotation(uint32_t k0, uint32_t k1, uint64_t serial, uint32_t step, double* m) {
  uint32_t a;
  double w = 1.0, x = 0.0, y = 0.0, z = 0.0;
  for (a = 0; a < 64; a++) {
    smolrng4 r = smolrng_mol(k0, k1, serial, step, SMOLRNG_PRODROT, a);
    double x1 = smolrng_sym(r.v[0]), y1 = smolrng_sym(r.v[1]);
    double x2 = smolrng_sym(r.v[2]), y2 = smolrng_sym(r.v[3]);
    double s1 = x1 * x1 + y1 * y1, s2 = x2 * x2 + y2 * y2;
    if (s1 < 1.0 && s2 > 0.0 && s2 < 1.0) {
      double f = sqrt((1.0 - s1) / s2);
      w = x1;
      x = y1;
      y = x2 * f;
      z = y2 * f;
      break;
    }
  }
  m[0] = 1.0 - 2.0 * (y * y + z * z);
  m[1] = 2.0 * (x * y - w * z);
  m[2] = 2.0 * (x * z + w * y);
  m[3] = 2.0 * (x * y + w * z);
  m[4] = 1.0 - 2.0 * (x * x + z * z);
  m[5] = 2.0 * (y * z - w * x);
  m[6] = 2.0 * (x * z - w * y);
  m[7] = 2.0 * (y * z + w * x);
  m[8] = 1.0 - 2.0 * (x * x + y * y);
}

#endif
