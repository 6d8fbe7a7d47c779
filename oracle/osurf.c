/* CPU oracle (test infrastructure only; never linked into the product).
 *
 * Surfaces of the oracle simulation: panels, surface actions, molecule-surface
 * collisions (checksurfaces), surface-bound actions (checksurfacebound) and the
 * surface test of bimolecular pairs (rxnXsurface), restated from
 * source/Smoldyn/smolsurface.c, the box-panel lists of smolboxes.c, and the
 * geometry helpers of source/libSteve/Geometry.c and SurfaceParam.c.
 *
 * Panel shapes: rectangle (dims 1-3), triangle (dims 1-3), sphere (dims 1-3),
 * cylinder, hemisphere and disk (dim 3).  Actions: reflect, transmit, absorb,
 * jump, and rate-based actions (adsorption with a given probability or rate,
 * desorption, state changes).  Surface-bound diffusion (movemol2closepanel),
 * 2D cylinders, hemispheres and disks, ports and emitters are not restated;
 * setting them up returns an error.
 *
 * Random draws: in SFMT mode the reference's global stream in program order;
 * in counter mode Philox draws keyed by (molecule serial, step, SMOLRNG_SURFACE,
 * pass<<8 | iteration) or, for pairs, (serial pair, step, SMOLRNG_RXNSURF). */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/smoldyn_amd_rng.h"
#include "oracle.h"
#include "osim.h"

#define OS_VERYCLOSE 1.0e-12 /* smoldyn.h:60 */
#define OS_VERYLARGE 1.0e+20 /* smoldyn.h:61 */

enum { OPS_RECT, OPS_TRI, OPS_SPH, OPS_CYL, OPS_HEMI, OPS_DISK, OPS_MAX };
enum { OPF_FRONT, OPF_BACK, OPF_NONE, OPF_BOTH };
enum {
  OSA_REFLECT, OSA_TRANS, OSA_ABSORB, OSA_JUMP, OSA_PORT, OSA_MULT, OSA_NO, OSA_NONE, OSA_ADSORB,
  OSA_REVDES, OSA_IRREVDES, OSA_FLIP
};
/* SurfParamAlgo values used here (SurfaceParam.h) */
enum { OSPA_IRRDES, OSPA_REVADS };

typedef struct oactdet { /* surfactionstruct, smoldyn.h */
  int newspec[O_MSMAX1];
  double rate[O_MSMAX1], prob[O_MSMAX1], cumprob[O_MSMAX1], revprob[O_MSMAX1];
  int datasrc[O_MSMAX1];
} oactdet;

typedef struct opanel {
  int ps, srf;
  double point[8][3];
  double front[3];
  int jumpp[2], jumpf[2];
} opanel;

typedef struct osurface {
  char name[O_NAMELEN];
  int action[O_MAXSPECIES][O_MSMAX][3];
  oactdet* det[O_MAXSPECIES][O_MSMAX][3];
  int npanel[OPS_MAX];
  int* panels[OPS_MAX]; /* panel ids in creation order, per shape */
} osurface;

#define OS_MAXSRF 16
typedef struct osurfss {
  int nsrf;
  osurface* srf[OS_MAXSRF];
  int npnl, maxpnl;
  opanel* pnl;
  double epsilon, margin, neighdist;
  int srfmollist[O_MAXLIST]; /* bit 1 diffuse, 2 react, 4 surface-bound */
  int* bpstart;              /* box panel lists: panels of box b are bplist[bpstart[b]..bpstart[b+1]) */
  int* bplist;
} osurfss;

enum { OSML_DIFFUSE = 1, OSML_REACT = 2, OSML_SRFBOUND = 4 };

/* ------------------------------------------------------------------------ */
/* random draws                                                              */

static double os_randCOD(osim* sim, const omol* mp, uint32_t sub) {
  if (sim->rngmode == ORNG_SFMT) return osim_randCOD(sim);
  smolrng4 r = smolrng_mol(sim->key0, sim->key1, mp->serno, sim->stepindex, SMOLRNG_SURFACE, sub);
  return smolrng_u01(r.v[0]);
}

static double os_randCCD(osim* sim, const omol* mp, uint32_t sub) {
  if (sim->rngmode == ORNG_SFMT) return osim_randCCD(sim);
  smolrng4 r = smolrng_mol(sim->key0, sim->key1, mp->serno, sim->stepindex, SMOLRNG_SURFACE, sub);
  return smolrng_u01cc(r.v[0]);
}

/* ------------------------------------------------------------------------ */
/* geometry (source/libSteve/Geometry.c)                                     */

/* Geo_SphereNormal (Geometry.c:199-213) */
double osurf_geo_spherenormal(const double* cent, const double* pt, int front, int dim, double* ans) {
  int d;
  double dist = 0;
  for (d = 0; d < dim; d++) {
    ans[d] = front * (pt[d] - cent[d]);
    dist += ans[d] * ans[d];
  }
  if (dist > 0) {
    dist = sqrt(dist);
    for (d = 0; d < dim; d++) ans[d] /= dist;
  } else {
    ans[0] = 1;
    for (d = 1; d < dim; d++) ans[d] = 0;
  }
  return dist;
}

/* Geo_PtInSlab (Geometry.c:495-505) */
int osurf_geo_ptinslab(const double* pt1, const double* pt2, const double* test, int dim) {
  double dot = 0;
  int d;
  for (d = 0; d < dim; d++) dot += (test[d] - pt1[d]) * (pt2[d] - pt1[d]);
  if (dot < 0) return 0;
  dot = 0;
  for (d = 0; d < dim; d++) dot += (test[d] - pt2[d]) * (pt1[d] - pt2[d]);
  if (dot < 0) return 0;
  return 1;
}

/* Geo_LineXSphs (Geometry.c:1200-1221) */
double osurf_geo_linexsphs(const double* pt1, const double* pt2, const double* cent, double rad, int dim,
                           double* crss2ptr, double* nrdistptr, double* nrposptr) {
  double a = 0, b = 0, c = 0, cross, cross2, radical, nrdist;
  int d;
  for (d = 0; d < dim; d++) {
    a += (pt2[d] - pt1[d]) * (pt2[d] - pt1[d]);
    b += (cent[d] - pt1[d]) * (pt2[d] - pt1[d]);
    c += (pt1[d] - cent[d]) * (pt1[d] - cent[d]);
  }
  nrdist = sqrt(c - b * b / a);
  if (nrdistptr) *nrdistptr = nrdist;
  if (nrposptr) *nrposptr = b / a;
  b *= -2;
  c -= rad * rad;
  radical = b * b - 4 * a * c;
  if (nrdist <= rad && radical < 0) radical = 0;
  radical = sqrt(radical);
  cross = (-b - radical) / (2 * a);
  cross2 = (-b + radical) / (2 * a);
  if (crss2ptr) *crss2ptr = cross2;
  return cross;
}

/* the eight corner dot products of Geo_RectXaabb3 / Geo_DiskXaabb3 */
static void os_corners(const double* v, const double* b1, const double* b2, double* o) {
  o[0] = v[0] * b1[0] + v[1] * b1[1] + v[2] * b1[2];
  o[1] = v[0] * b1[0] + v[1] * b1[1] + v[2] * b2[2];
  o[2] = v[0] * b1[0] + v[1] * b2[1] + v[2] * b1[2];
  o[3] = v[0] * b1[0] + v[1] * b2[1] + v[2] * b2[2];
  o[4] = v[0] * b2[0] + v[1] * b1[1] + v[2] * b1[2];
  o[5] = v[0] * b2[0] + v[1] * b1[1] + v[2] * b2[2];
  o[6] = v[0] * b2[0] + v[1] * b2[1] + v[2] * b1[2];
  o[7] = v[0] * b2[0] + v[1] * b2[1] + v[2] * b2[2];
}
static int os_all_lt(const double* o, double c) {
  int k;
  for (k = 0; k < 8; k++)
    if (!(o[k] < c)) return 0;
  return 1;
}
static int os_all_gt(const double* o, double c) {
  int k;
  for (k = 0; k < 8; k++)
    if (!(o[k] > c)) return 0;
  return 1;
}

/* Geo_RectXaabb3 (Geometry.c:1640-1700): separating-axis test of the
   parallelogram r1,r2,r3(,r4) against the box [bpt1,bpt2]. */
int osurf_geo_rectxaabb3(const double* r1, const double* r2, const double* r3, const double* r4,
                         const double* bpt1, const double* bpt2) {
  double r5[3], r6[3], r7[3], v1[3], o[8], cmp;
  const double* rr[7];
  int d, k, s;
  for (d = 0; d < 3; d++) {
    r5[d] = r2[d] + r4[d] - r1[d];
    r6[d] = r3[d] + r4[d] - r1[d];
    r7[d] = r2[d] + r3[d] - r1[d];
  }
  rr[0] = r1;
  rr[1] = r2;
  rr[2] = r3;
  rr[3] = r4;
  rr[4] = r5;
  rr[5] = r6;
  rr[6] = r7;
  for (d = 0; d < 3; d++) {
    int lo = 1, hi = 1;
    for (k = 0; k < 7; k++) {
      if (!(rr[k][d] < bpt1[d])) lo = 0;
      if (!(rr[k][d] > bpt2[d])) hi = 0;
    }
    if (lo || hi) return 0;
  }
  for (s = 0; s < 3; s++) {
    const double* ra = (s == 0) ? r2 : (s == 1) ? r3 : r4;
    v1[0] = ra[0] - r1[0];
    v1[1] = ra[1] - r1[1];
    v1[2] = ra[2] - r1[2];
    cmp = v1[0] * r1[0] + v1[1] * r1[1] + v1[2] * r1[2];
    os_corners(v1, bpt1, bpt2, o);
    if (os_all_lt(o, cmp)) return 0;
    cmp = v1[0] * ra[0] + v1[1] * ra[1] + v1[2] * ra[2];
    if (os_all_gt(o, cmp)) return 0;
  }
  return 1;
}

/* Geo_SphsXaabb3 (Geometry.c:1736-1775); every path after the projection
   tests returns 1 except "box entirely inside the sphere". */
int osurf_geo_sphsxaabb3(const double* cent, double rad, const double* bpt1, const double* bpt2) {
  double mn, mx, d2, r2;
  int k;
  if (cent[0] + rad < bpt1[0]) return 0;
  if (cent[0] - rad > bpt2[0]) return 0;
  if (cent[1] + rad < bpt1[1]) return 0;
  if (cent[1] - rad > bpt2[1]) return 0;
  if (cent[2] + rad < bpt1[2]) return 0;
  if (cent[2] - rad > bpt2[2]) return 0;
  r2 = rad * rad;
  mn = mx = 0;
  for (k = 0; k < 8; k++) {
    const double* bx = (k & 4) ? bpt2 : bpt1;
    const double* by = (k & 2) ? bpt2 : bpt1;
    const double* bz = (k & 1) ? bpt2 : bpt1;
    d2 = (bx[0] - cent[0]) * (bx[0] - cent[0]) + (by[1] - cent[1]) * (by[1] - cent[1]) +
         (bz[2] - cent[2]) * (bz[2] - cent[2]);
    if (k == 0) mn = mx = d2;
    else if (d2 < mn) mn = d2;
    else if (d2 > mx) mx = d2;
  }
  if (r2 > mx) return 0;
  return 1;
}

/* Geo_DiskXaabb3 (Geometry.c:1918-1942) */
int osurf_geo_diskxaabb3(const double* cent, double rad, const double* norm, const double* bpt1,
                         const double* bpt2) {
  double dot, cmp, o[8];
  dot = sqrt(norm[1] * norm[1] + norm[2] * norm[2]);
  if (cent[0] - rad * dot > bpt2[0]) return 0;
  if (cent[0] + rad * dot < bpt1[0]) return 0;
  dot = sqrt(norm[0] * norm[0] + norm[2] * norm[2]);
  if (cent[1] - rad * dot > bpt2[1]) return 0;
  if (cent[1] + rad * dot < bpt1[1]) return 0;
  dot = sqrt(norm[0] * norm[0] + norm[1] * norm[1]);
  if (cent[2] - rad * dot > bpt2[2]) return 0;
  if (cent[2] + rad * dot < bpt1[2]) return 0;
  cmp = norm[0] * cent[0] + norm[1] * cent[1] + norm[2] * cent[2];
  os_corners(norm, bpt1, bpt2, o);
  if (os_all_lt(o, cmp)) return 0;
  if (os_all_gt(o, cmp)) return 0;
  return 1;
}

/* Geo_LineXaabb2 (Geometry.c:1544-1558) */
int osurf_geo_linexaabb2(const double* pt1, const double* pt2, const double* norm, const double* bpt1,
                         const double* bpt2) {
  double cmp, b1, b2, b3, b4;
  if (pt1[0] < bpt1[0] && pt2[0] < bpt1[0]) return 0;
  if (pt1[0] > bpt2[0] && pt2[0] > bpt2[0]) return 0;
  if (pt1[1] < bpt1[1] && pt2[1] < bpt1[1]) return 0;
  if (pt1[1] > bpt2[1] && pt2[1] > bpt2[1]) return 0;
  cmp = norm[0] * pt1[0] + norm[1] * pt1[1];
  b1 = norm[0] * bpt1[0] + norm[1] * bpt1[1];
  b2 = norm[0] * bpt1[0] + norm[1] * bpt2[1];
  b3 = norm[0] * bpt2[0] + norm[1] * bpt1[1];
  b4 = norm[0] * bpt2[0] + norm[1] * bpt2[1];
  if (b1 < cmp && b2 < cmp && b3 < cmp && b4 < cmp) return 0;
  if (b1 > cmp && b2 > cmp && b3 > cmp && b4 > cmp) return 0;
  return 1;
}

/* Geo_CircleXaabb2 (Geometry.c:1705-1733) */
int osurf_geo_circlexaabb2(const double* cent, double rad, const double* bpt1, const double* bpt2) {
  double mn, mx, d2, r2;
  int k;
  if (cent[0] + rad < bpt1[0]) return 0;
  if (cent[0] - rad > bpt2[0]) return 0;
  if (cent[1] + rad < bpt1[1]) return 0;
  if (cent[1] - rad > bpt2[1]) return 0;
  r2 = rad * rad;
  mn = mx = 0;
  for (k = 0; k < 4; k++) {
    const double* bx = (k & 2) ? bpt2 : bpt1;
    const double* by = (k & 1) ? bpt2 : bpt1;
    d2 = (bx[0] - cent[0]) * (bx[0] - cent[0]) + (by[1] - cent[1]) * (by[1] - cent[1]);
    if (k == 0) mn = mx = d2;
    else if (d2 < mn) mn = d2;
    else if (d2 > mx) mx = d2;
  }
  if (r2 > mx) return 0;
  if (r2 >= mn) return 1;
  if (cent[0] >= bpt1[0] && cent[0] <= bpt2[0]) return 1;
  if (cent[1] >= bpt1[1] && cent[1] <= bpt2[1]) return 1;
  return 0;
}

/* Geo_LineNormal (Geometry.c:58-71): unit normal (dy, -dx) of the line
   pt1 -> pt2, scaled by the reciprocal of its length */
double osurf_geo_linenormal(const double* pt1, const double* pt2, double* ans) {
  double dx = pt2[0] - pt1[0], dy = pt2[1] - pt1[1], len = sqrt(dx * dx + dy * dy), leninv;
  if (len > 0) {
    leninv = 1.0 / len;
    ans[0] = leninv * dy;
    ans[1] = -leninv * dx;
  } else {
    ans[0] = 1;
    ans[1] = 0;
  }
  return len;
}

/* Geo_TriNormal (Geometry.c:176-195): unit normal (pt2-pt1) x (pt3-pt2);
   a triangle of area below 50 DBL_EPSILON takes the normal of its first edge
   in the xy plane.  Returns the area. */
double osurf_geo_trinormal(const double* pt1, const double* pt2, const double* pt3, double* ans) {
  double e1[3], e2[3], len;
  int d;
  for (d = 0; d < 3; d++) {
    e1[d] = pt2[d] - pt1[d];
    e2[d] = pt3[d] - pt2[d];
  }
  ans[0] = e1[1] * e2[2] - e1[2] * e2[1];
  ans[1] = -e1[0] * e2[2] + e1[2] * e2[0];
  ans[2] = e1[0] * e2[1] - e1[1] * e2[0];
  len = sqrt(ans[0] * ans[0] + ans[1] * ans[1] + ans[2] * ans[2]);
  if (len > 100 * DBL_EPSILON) {
    for (d = 0; d < 3; d++) ans[d] /= len;
  } else {
    osurf_geo_linenormal(pt1, pt2, ans);
    ans[2] = 0;
  }
  return len / 2;
}

/* Geo_UnitCross (Geometry.c:219-247): unit vector of (v1end - v1start) x
   (v2end - v2start), a NULL start meaning the origin; the zero vector when
   the cross product's length is at most 100 DBL_EPSILON.  Returns the length. */
double osurf_geo_unitcross(const double* v1start, const double* v1end, const double* v2start, const double* v2end,
                           double* ans) {
  double a[3], b[3], c[3], len;
  int d;
  for (d = 0; d < 3; d++) {
    a[d] = v1start ? v1end[d] - v1start[d] : v1end[d];
    b[d] = v2start ? v2end[d] - v2start[d] : v2end[d];
  }
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  if (len > 100 * DBL_EPSILON) {
    for (d = 0; d < 3; d++) ans[d] = c[d] / len;
    return len;
  }
  ans[0] = ans[1] = ans[2] = 0;
  return 0;
}

/* Geo_PtInTriangle2 (Geometry.c:474-492): test is in the triangle (vertices
   point[0..2], outward edge normals point[3..5]) when it is on no edge's
   outer side */
int osurf_geo_ptintriangle2(const double (*point)[3], const double* test) {
  int k;
  for (k = 0; k < 3; k++) {
    const double* pt = point[k];
    const double* en = point[k + 3];
    double dot = (test[0] - pt[0]) * en[0] + (test[1] - pt[1]) * en[1] + (test[2] - pt[2]) * en[2];
    if (dot > 0) return 0;
  }
  return 1;
}

/* Geo_TriXaabb3 (Geometry.c:1585-1605): 0 when the triangle's bounding box
   misses the box bpt1-bpt2 or all 8 box corners lie strictly on one side of
   the triangle's plane, else 1 */
int osurf_geo_trixaabb3(const double* pt1, const double* pt2, const double* pt3, const double* norm,
                        const double* bpt1, const double* bpt2) {
  double cmp, o[8];
  int d;
  for (d = 0; d < 3; d++) {
    if (pt1[d] < bpt1[d] && pt2[d] < bpt1[d] && pt3[d] < bpt1[d]) return 0;
    if (pt1[d] > bpt2[d] && pt2[d] > bpt2[d] && pt3[d] > bpt2[d]) return 0;
  }
  cmp = norm[0] * pt1[0] + norm[1] * pt1[1] + norm[2] * pt1[2];
  os_corners(norm, bpt1, bpt2, o);
  if (os_all_lt(o, cmp)) return 0;
  if (os_all_gt(o, cmp)) return 0;
  return 1;
}

/* Geo_LineNormal3D (Geometry.c:103-160): unit vector from the line pt1-pt2
   to point, perpendicular to the line, and the distance.  Degenerate cases:
   pt1 == pt2 gives the unit vector to point (unit x when point == pt1); a
   point on the line gives (line_y, -line_x, 0) normalized, or unit x for a
   line along z, and distance 0. */
double osurf_geo_linenormal3d(const double* pt1, const double* pt2, const double* point, double* ans) {
  const double eps = 100 * DBL_EPSILON;
  double line[3], dot;
  int d;
  for (d = 0; d < 3; d++) line[d] = pt2[d] - pt1[d];
  dot = line[0] * line[0] + line[1] * line[1] + line[2] * line[2];
  if (dot < eps) {
    for (d = 0; d < 3; d++) ans[d] = point[d] - pt1[d];
    dot = ans[0] * ans[0] + ans[1] * ans[1] + ans[2] * ans[2];
    if (dot < eps) {
      ans[0] = 1;
      ans[1] = ans[2] = 0;
      return 0;
    }
    dot = sqrt(dot);
    for (d = 0; d < 3; d++) ans[d] /= dot;
    return dot;
  }
  dot = sqrt(dot);
  for (d = 0; d < 3; d++) line[d] /= dot;
  for (d = 0; d < 3; d++) ans[d] = point[d] - pt1[d];
  for (int pass = 0; pass < 2; pass++) { /* projection removed twice */
    dot = ans[0] * line[0] + ans[1] * line[1] + ans[2] * line[2];
    for (d = 0; d < 3; d++) ans[d] -= dot * line[d];
  }
  dot = ans[0] * ans[0] + ans[1] * ans[1] + ans[2] * ans[2];
  if (dot < eps) {
    if (line[0] == 0 && line[1] == 0) {
      ans[0] = 1;
      ans[1] = ans[2] = 0;
      return 0;
    }
    ans[0] = line[1];
    ans[1] = -line[0];
    ans[2] = 0;
    dot = sqrt(ans[0] * ans[0] + ans[1] * ans[1] + ans[2] * ans[2]);
    for (d = 0; d < 3; d++) ans[d] /= dot;
    return 0;
  }
  dot = sqrt(dot);
  for (d = 0; d < 3; d++) ans[d] /= dot;
  return dot;
}

/* Geo_LineXCyls (Geometry.c:1248-1258): crossings of the segment pt1-pt2
   with the infinite cylinder of axis cp1-cp2 and radius rad, found as the
   crossings of the segment's perpendicular offsets from the axis with a
   circle about the origin */
double osurf_geo_linexcyls(const double* pt1, const double* pt2, const double* cp1, const double* cp2, double rad,
                           double* crss2ptr, double* nrdistptr, double* nrposptr) {
  double r1, r2, n1[3], n2[3], zero[3] = {0, 0, 0};
  int d;
  r1 = osurf_geo_linenormal3d(cp1, cp2, pt1, n1);
  r2 = osurf_geo_linenormal3d(cp1, cp2, pt2, n2);
  for (d = 0; d < 3; d++) {
    n1[d] *= r1;
    n2[d] *= r2;
  }
  return osurf_geo_linexsphs(n1, n2, zero, rad, 3, crss2ptr, nrdistptr, nrposptr);
}

/* Geo_LineXaabb (Geometry.c:1561-1582): slab test of the segment (or, with
   infline, the whole line) through pt1, pt2 against the box bpt1-bpt2 */
int osurf_geo_linexaabb(const double* pt1, const double* pt2, const double* bpt1, const double* bpt2, int dim,
                        int infline) {
  double nearhi = -1.0, farlo = 2.0, denom, a1, a2;
  int d;
  for (d = 0; d < dim; d++) {
    denom = pt2[d] - pt1[d];
    if (denom == 0) {
      if (pt1[d] < bpt1[d] || pt1[d] > bpt2[d]) return 0;
    } else {
      a1 = (bpt1[d] - pt1[d]) / denom;
      a2 = (bpt2[d] - pt1[d]) / denom;
      if (a1 < a2) {
        if (a1 > nearhi) nearhi = a1;
        if (a2 < farlo) farlo = a2;
      } else {
        if (a2 > nearhi) nearhi = a2;
        if (a1 < farlo) farlo = a1;
      }
    }
  }
  if (infline) return nearhi <= farlo;
  return nearhi <= farlo && nearhi <= 1 && farlo >= 0;
}

/* Geo_CylisXaabb3 (Geometry.c:1777-1916): does the infinite cylinder of axis
   pt1-pt2 and radius rad meet the box?  The box corners are projected onto
   the plane perpendicular to the axis (axes v1, v2); separating-axis tests
   along the three projected box edges, then the corner distances to the
   circle, the axis line against the box, and the 12 projected box edges
   against the circle. */
int osurf_geo_cylisxaabb3(const double* pt1, const double* pt2, double rad, const double* bpt1, const double* bpt2) {
  /* projected box edges of test 5, as corner index pairs (corner k takes
     bpt2 in x, y, z for bits 4, 2, 1) */
  static const int edge[12][2] = {{0, 1}, {1, 3}, {3, 2}, {2, 0}, {0, 4}, {1, 5},
                                  {2, 6}, {3, 7}, {4, 5}, {5, 7}, {7, 6}, {6, 4}};
  static const int axis[3] = {1, 2, 4};
  double v1[3], v2[3], v3[3], cent[2], ab[8][2], o1[8], o2[8], nrm, b[8], cmp, r2, d2, mn, mx, crss1, crss2;
  int k, a;
  osurf_geo_linenormal3d(pt1, pt2, pt1, v1);
  v3[0] = pt2[0] - pt1[0];
  v3[1] = pt2[1] - pt1[1];
  v3[2] = pt2[2] - pt1[2];
  nrm = sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
  if (!nrm) return 0;
  for (k = 0; k < 3; k++) v3[k] /= nrm;
  v2[0] = v3[1] * v1[2] - v3[2] * v1[1];
  v2[1] = v3[2] * v1[0] - v3[0] * v1[2];
  v2[2] = v3[0] * v1[1] - v3[1] * v1[0];
  cent[0] = v1[0] * pt1[0] + v1[1] * pt1[1] + v1[2] * pt1[2];
  cent[1] = v2[0] * pt1[0] + v2[1] * pt1[1] + v2[2] * pt1[2];
  os_corners(v1, bpt1, bpt2, o1);
  os_corners(v2, bpt1, bpt2, o2);
  for (k = 0; k < 8; k++) {
    ab[k][0] = o1[k];
    ab[k][1] = o2[k];
  }
  for (a = 0; a < 3; a++) { /* test 1 */
    double u[2];
    u[0] = ab[axis[a]][0] - ab[0][0];
    u[1] = ab[axis[a]][1] - ab[0][1];
    nrm = sqrt(u[0] * u[0] + u[1] * u[1]);
    if (nrm > 0) {
      u[0] /= nrm;
      u[1] /= nrm;
      for (k = 0; k < 8; k++) b[k] = u[0] * ab[k][0] + u[1] * ab[k][1];
      cmp = u[0] * cent[0] + u[1] * cent[1] - rad;
      if (os_all_lt(b, cmp)) return 0;
      cmp = u[0] * cent[0] + u[1] * cent[1] + rad;
      if (os_all_gt(b, cmp)) return 0;
    }
  }
  r2 = rad * rad; /* tests 2, 3 */
  mn = mx = 0;
  for (k = 0; k < 8; k++) {
    d2 = (ab[k][0] - cent[0]) * (ab[k][0] - cent[0]) + (ab[k][1] - cent[1]) * (ab[k][1] - cent[1]);
    if (k == 0) mn = mx = d2;
    else if (d2 < mn) mn = d2;
    else if (d2 > mx) mx = d2;
  }
  if (r2 > mx) return 0;
  if (r2 >= mn) return 1;
  if (osurf_geo_linexaabb(pt1, pt2, bpt1, bpt2, 3, 1)) return 1; /* test 4 */
  for (k = 0; k < 12; k++) {                                    /* test 5 */
    crss1 = osurf_geo_linexsphs(ab[edge[k][0]], ab[edge[k][1]], cent, rad, 2, &crss2, NULL, NULL);
    if ((crss1 >= 0 && crss1 <= 1) || (crss2 >= 0 && crss2 <= 1)) return 1;
  }
  return 0;
}

/* Geo_CylsXaabb3 (Geometry.c:1964-1970) with Geo_Cyl2Rect (:1137-1162): the
   infinite cylinder test, then the rectangle through the axis spanned by the
   axis and the cylinder's diameter along v2 = v1 x axis */
int osurf_geo_cylsxaabb3(const double* pt1, const double* pt2, double rad, const double* bpt1, const double* bpt2) {
  double v1[3], v2[3], nrm, r[4][3];
  int d;
  if (!osurf_geo_cylisxaabb3(pt1, pt2, rad, bpt1, bpt2)) return 0;
  osurf_geo_linenormal3d(pt1, pt2, pt1, v1);
  v2[0] = v1[1] * (pt2[2] - pt1[2]) - v1[2] * (pt2[1] - pt1[1]);
  v2[1] = v1[2] * (pt2[0] - pt1[0]) - v1[0] * (pt2[2] - pt1[2]);
  v2[2] = v1[0] * (pt2[1] - pt1[1]) - v1[1] * (pt2[0] - pt1[0]);
  nrm = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  for (d = 0; d < 3; d++) v2[d] /= nrm;
  for (d = 0; d < 3; d++) {
    r[0][d] = pt1[d] + rad * (-v1[d] - v2[d]);
    r[1][d] = pt1[d] + rad * (+v1[d] - v2[d]);
    r[2][d] = pt1[d] + rad * (-v1[d] + v2[d]);
    r[3][d] = pt2[d] + rad * (-v1[d] - v2[d]);
  }
  return osurf_geo_rectxaabb3(r[0], r[1], r[2], r[3], bpt1, bpt2);
}

/* Geo_HemisXaabb3 (Geometry.c:1956-1962) with Geo_Hemis2Rect (:1113-1135):
   the sphere test, then the box of side 2 rad x 2 rad x rad behind the
   hemisphere's opening (outvect points away from the dome) */
int osurf_geo_hemisxaabb3(const double* cent, double rad, const double* outvect, const double* bpt1,
                          const double* bpt2) {
  double v1[3], v2[3], zero[3] = {0, 0, 0}, r[4][3];
  int d;
  if (!osurf_geo_sphsxaabb3(cent, rad, bpt1, bpt2)) return 0;
  osurf_geo_linenormal3d(zero, outvect, zero, v1);
  v2[0] = v1[1] * outvect[2] - v1[2] * outvect[1];
  v2[1] = v1[2] * outvect[0] - v1[0] * outvect[2];
  v2[2] = v1[0] * outvect[1] - v1[1] * outvect[0];
  for (d = 0; d < 3; d++) {
    r[0][d] = cent[d] + rad * (-v1[d] - v2[d]);
    r[1][d] = cent[d] + rad * (+v1[d] - v2[d]);
    r[2][d] = cent[d] + rad * (-v1[d] + v2[d]);
    r[3][d] = cent[d] + rad * (-v1[d] - v2[d] - outvect[d]);
  }
  return osurf_geo_rectxaabb3(r[0], r[1], r[2], r[3], bpt1, bpt2);
}

/* desorbdist (SurfaceParam.c:202-220) for the two algorithms desorption uses */
double osurf_desorbdist_y(double step, double y, int revads) {
  double y2 = y * y, x;
  if (!revads) x = (0.571825 * y - 0.552246 * y2) / (1.0 - 1.53908 * y + 0.546424 * y2);
  else x = (0.729614 * y - 0.70252 * y2) / (1.0 - 1.47494 * y + 0.484371 * y2);
  return step * x;
}

/* ------------------------------------------------------------------------ */
/* setup                                                                     */

static osurfss* os_ss(osim* sim) {
  if (!sim->srfss) {
    osurfss* ss = (osurfss*)calloc(1, sizeof(osurfss));
    ss->epsilon = 100 * DBL_EPSILON; /* surfenablesurfaces, smolsurface.c:1235-1237 */
    ss->margin = 100 * DBL_EPSILON;
    ss->neighdist = 10 * DBL_EPSILON;
    sim->srfss = ss;
  }
  return sim->srfss;
}

void osurf_free(osim* sim) {
  osurfss* ss = sim->srfss;
  int s, ps, i, ms, f;
  if (!ss) return;
  for (s = 0; s < ss->nsrf; s++) {
    osurface* srf = ss->srf[s];
    for (ps = 0; ps < OPS_MAX; ps++) free(srf->panels[ps]);
    for (i = 0; i < O_MAXSPECIES; i++)
      for (ms = 0; ms < O_MSMAX; ms++)
        for (f = 0; f < 3; f++) free(srf->det[i][ms][f]);
    free(srf);
  }
  free(ss->pnl);
  free(ss->bpstart);
  free(ss->bplist);
  free(ss);
  sim->srfss = NULL;
}

/* surfaddsurface: default actions transmit on both faces, none for PFnone
   (surfacealloc, smolsurface.c:1110-1118). */
int osim_add_surface(osim* sim, const char* name) {
  osurfss* ss = os_ss(sim);
  osurface* srf;
  int i, ms;
  if (ss->nsrf == OS_MAXSRF) return -1;
  srf = (osurface*)calloc(1, sizeof(osurface));
  strncpy(srf->name, name, O_NAMELEN - 1);
  for (i = 0; i < O_MAXSPECIES; i++)
    for (ms = 0; ms < O_MSMAX; ms++) {
      srf->action[i][ms][OPF_FRONT] = srf->action[i][ms][OPF_BACK] = OSA_TRANS;
      srf->action[i][ms][OPF_NONE] = OSA_NO;
    }
  ss->srf[ss->nsrf] = srf;
  sim->ok = 0;
  return ss->nsrf++;
}

static oactdet* os_actalloc(int species) { /* surfaceactionalloc, smolsurface.c:786-826 */
  oactdet* a = (oactdet*)calloc(1, sizeof(oactdet));
  int ms;
  for (ms = 0; ms < O_MSMAX1; ms++) a->newspec[ms] = species;
  return a;
}

/* surfsetaction (smolsurface.c:2158-2276) for one species (ident>0) or all
   (ident==-5); newident>0 records a species change for the action. */
int osim_set_surface_action(osim* sim, int s, int ident, int ms, int face, int act, int newident) {
  osurfss* ss = sim->srfss;
  osurface* srf;
  int i, ms1, mslo, mshi, ilo, ihi;
  if (!ss || s < 0 || s >= ss->nsrf) return 1;
  srf = ss->srf[s];
  if (ms == OMS_BSOLN || ms == OMS_NONE) return 2;
  if (ms == OMS_ALL) {
    mslo = 0;
    mshi = O_MSMAX - 1;
  } else mslo = mshi = ms;
  if (ident > 0) ilo = ihi = ident;
  else if (ident == -5) {
    ilo = 1;
    ihi = sim->nspecies - 1;
  } else return 3;
  if (face == OPF_FRONT || face == OPF_BACK || face == OPF_BOTH) {
    if (!(act == OSA_REFLECT || act == OSA_TRANS || act == OSA_ABSORB || act == OSA_JUMP || act == OSA_MULT))
      return 3;
  } else if (!(act == OSA_MULT || act == OSA_NO)) return 3;
  for (i = ilo; i <= ihi; i++)
    for (ms1 = mslo; ms1 <= mshi; ms1++) {
      if (face == OPF_NONE) srf->action[i][ms1][OPF_NONE] = act;
      if (face == OPF_FRONT || face == OPF_BOTH) srf->action[i][ms1][OPF_FRONT] = act;
      if (face == OPF_BACK || face == OPF_BOTH) srf->action[i][ms1][OPF_BACK] = act;
    }
  if (newident > 0 && ident > 0) {
    i = ident;
    for (ms1 = mslo; ms1 <= mshi; ms1++) {
      int f;
      for (f = 0; f < 2; f++) {
        if (!(face == OPF_BOTH || face == f)) continue;
        if (!srf->det[i][ms1][f]) srf->det[i][ms1][f] = os_actalloc(i);
        srf->det[i][ms1][f]->datasrc[ms1] = 3;
        srf->det[i][ms1][f]->newspec[ms1] = newident;
      }
    }
  } else if (newident > 0 && ident == -5) {
    for (i = 1; i < sim->nspecies; i++)
      for (ms1 = mslo; ms1 <= mshi; ms1++) {
        int f;
        for (f = 0; f < 2; f++) {
          if (!(face == OPF_BOTH || face == f)) continue;
          if (!srf->det[i][ms1][f]) srf->det[i][ms1][f] = os_actalloc(i);
          srf->det[i][ms1][f]->datasrc[ms1] = 3;
          srf->det[i][ms1][f]->newspec[ms1] = newident;
        }
      }
  }
  sim->ok = 0;
  return 0;
}

/* srftristate2index (smolsurface.c:713-740) */
static void os_tristate2index(int ms, int ms1, int ms2, int* ms3p, int* facep, int* ms4p) {
  int ms3, face, ms4;
  if (ms == OMS_NONE) ms = OMS_SOLN;
  if (ms == OMS_SOLN && (ms1 == OMS_SOLN || ms1 == OMS_BSOLN)) ms3 = OMS_SOLN;
  else if (ms == OMS_SOLN) ms3 = ms1;
  else ms3 = ms;
  if (ms1 == OMS_SOLN) face = OPF_FRONT;
  else if (ms1 == OMS_BSOLN) face = OPF_BACK;
  else face = OPF_NONE;
  ms4 = ms2;
  if (ms != OMS_SOLN && ms1 != OMS_SOLN && ms1 != OMS_BSOLN && ms1 != ms) {
    ms3 = OMS_NONE;
    face = OPF_NONE;
    ms4 = OMS_NONE;
  }
  *ms3p = ms3;
  *facep = face;
  *ms4p = ms4;
}

/* surfsetrate (smolsurface.c:2278-2329): which 1 = rate, 2 = probability
   (rate_internal); newident -5 keeps the species. */
int osim_set_surface_rate(osim* sim, int s, int ident, int ms, int ms1, int ms2, int newident, double value,
                          int which) {
  osurfss* ss = sim->srfss;
  osurface* srf;
  int ms3, face, ms4, i, ilo, ihi;
  if (!ss || s < 0 || s >= ss->nsrf) return 1;
  srf = ss->srf[s];
  if (ms == OMS_BSOLN || ms == OMS_ALL) return 2;
  if (ms1 > OMS_BSOLN) return 3;
  if (ms != OMS_SOLN && !(ms1 == OMS_SOLN || ms1 == OMS_BSOLN || ms1 == ms)) return 3;
  if (ms2 > OMS_BSOLN || ms2 == ms1) return 4;
  if ((newident != -5 && newident < 0) || newident >= O_MAXSPECIES) return 5;
  if (value < 0 && !(which == 1 && value == -1)) return 6;
  if (which == 2 && value > 1) return 6;
  os_tristate2index(ms, ms1, ms2, &ms3, &face, &ms4);
  if (ms3 == OMS_NONE) return 3;
  if (ident > 0) ilo = ihi = ident;
  else if (ident == -5) {
    ilo = 1;
    ihi = sim->nspecies - 1;
  } else return 1;
  for (i = ilo; i <= ihi; i++) {
    oactdet* a;
    if (!srf->det[i][ms3][face]) srf->det[i][ms3][face] = os_actalloc(i);
    a = srf->det[i][ms3][face];
    srf->action[i][ms3][face] = OSA_MULT;
    if (which == 1) {
      a->rate[ms4] = value;
      a->datasrc[ms4] = 1;
    } else {
      a->prob[ms4] = value;
      a->datasrc[ms4] = 2;
    }
    a->newspec[ms4] = (newident == -5) ? i : newident;
  }
  sim->ok = 0;
  return 0;
}

/* surfaddpanel (smolsurface.c:2432-2640) for rect, tri, sph (all dims), cyl,
   hemi, disk (dim 3)
   (dim 3).  rectaxis is the rectangle's "+0"/"-1"... string.  Returns the
   panel id or <0. */
int osim_add_panel(osim* sim, int s, int ps, const char* rectaxis, const double* params) {
  osurfss* ss = sim->srfss;
  osurface* srf;
  opanel* pnl;
  int dim = sim->dim, pt, pdim, d, axis01, axis12;
  double(*point)[3];
  double* front;
  if (!ss || s < 0 || s >= ss->nsrf) return -1;
  srf = ss->srf[s];
  if (ss->npnl == ss->maxpnl) {
    ss->maxpnl = 2 * ss->maxpnl + 16;
    ss->pnl = (opanel*)realloc(ss->pnl, ss->maxpnl * sizeof(opanel));
  }
  pnl = &ss->pnl[ss->npnl];
  memset(pnl, 0, sizeof(opanel));
  pnl->ps = ps;
  pnl->srf = s;
  pnl->jumpp[0] = pnl->jumpp[1] = -1;
  pnl->jumpf[0] = pnl->jumpf[1] = OPF_NONE;
  point = pnl->point;
  front = pnl->front;
  if (ps == OPS_RECT) {
    char ch;
    if (!rectaxis) return -3;
    ch = rectaxis[0];
    if (!(ch == '+' || ch == '-')) return -3;
    if (rectaxis[1] == '0' || rectaxis[1] == 'x') pdim = 0;
    else if (rectaxis[1] == '1' || rectaxis[1] == 'y') pdim = 1;
    else if (rectaxis[1] == '2' || rectaxis[1] == 'z') pdim = 2;
    else pdim = 4;
    if (pdim >= dim) return -3;
    if (dim == 1) {
      point[0][0] = params[0];
      front[0] = (double)(ch == '+' ? 1 : -1);
      front[1] = 0;
    } else if (dim == 2) {
      pt = 1;
      if (ch == '-') pt *= -1;
      if (pdim == 0) pt *= -1;
      if (params[2] < 0) pt *= -1;
      if (pt < 0) pt = 0;
      point[pt][0] = params[0];
      point[pt][1] = params[1];
      point[!pt][0] = params[0] + (pdim == 0 ? 0 : params[2]);
      point[!pt][1] = params[1] + (pdim == 1 ? 0 : params[2]);
      front[0] = (double)(ch == '+' ? 1 : -1);
      front[1] = (double)pdim;
      front[2] = (double)(!pdim);
      point[2][0] = front[0] * front[1];
      point[2][1] = -front[0] * front[2];
      point[3][0] = -front[0] * front[1];
      point[3][1] = front[0] * front[2];
    } else {
      pt = 1;
      if (ch == '-') pt *= -1;
      if (params[3] * params[4] < 0) pt *= -1;
      if (pt < 0) pt = 3;
      point[0][0] = params[0];
      point[0][1] = params[1];
      point[0][2] = params[2];
      point[pt][0] = params[0] + (pdim == 2 ? params[3] : 0);
      point[pt][1] = params[1] + (pdim == 0 ? params[3] : 0);
      point[pt][2] = params[2] + (pdim == 1 ? params[4] : 0);
      point[2][0] = params[0] + (pdim == 0 ? 0 : params[3]);
      point[2][1] = params[1] + (pdim == 1 ? 0 : (pdim == 0 ? params[3] : params[4]));
      point[2][2] = params[2] + (pdim == 2 ? 0 : params[4]);
      point[4 - pt][0] = params[0] + (pdim == 1 ? params[3] : 0);
      point[4 - pt][1] = params[1] + (pdim == 2 ? params[4] : 0);
      point[4 - pt][2] = params[2] + (pdim == 0 ? params[4] : 0);
      front[0] = (double)(ch == '+' ? 1 : -1);
      front[1] = (double)pdim;
      if (pdim == 0) axis01 = (pt == 1 ? 1 : 2);
      else if (pdim == 1) axis01 = (pt == 1 ? 2 : 0);
      else axis01 = (pt == 1 ? 0 : 1);
      axis12 = (axis01 + 1) % 3;
      if (axis12 == pdim) axis12 = (axis12 + 1) % 3;
      front[2] = (double)axis01;
      for (d = 0; d < 3; d++) point[4][d] = point[5][d] = point[6][d] = point[7][d] = 0;
      point[4][axis12] = (point[2][axis12] > point[1][axis12] ? -1 : 1);
      point[5][axis01] = (point[1][axis01] > point[0][axis01] ? 1 : -1);
      point[6][axis12] = (point[2][axis12] > point[1][axis12] ? 1 : -1);
      point[7][axis01] = (point[1][axis01] > point[0][axis01] ? -1 : 1);
    }
  } else if (ps == OPS_TRI) {
    if (dim == 1) {
      point[0][0] = params[0];
      front[0] = (double)1;
    } else if (dim == 2) {
      point[0][0] = params[0];
      point[0][1] = params[1];
      point[1][0] = params[2];
      point[1][1] = params[3];
      osurf_geo_linenormal(point[0], point[1], front);
      point[2][0] = front[1];
      point[2][1] = -front[0];
      point[3][0] = -front[1];
      point[3][1] = front[0];
    } else {
      /* vertices, the unit normal, then the outward unit normal of each edge
         (smolsurface.c:2521-2534) */
      for (pt = 0; pt < 3; pt++)
        for (d = 0; d < 3; d++) point[pt][d] = params[3 * pt + d];
      osurf_geo_trinormal(point[0], point[1], point[2], front);
      osurf_geo_unitcross(point[0], point[1], NULL, front, point[3]);
      osurf_geo_unitcross(point[1], point[2], NULL, front, point[4]);
      osurf_geo_unitcross(point[2], point[0], NULL, front, point[5]);
    }
  } else if (ps == OPS_SPH) {
    double sg;
    if (dim == 1) {
      point[0][0] = params[0];
      point[1][0] = fabs(params[1]);
      sg = params[1];
    } else if (dim == 2) {
      point[0][0] = params[0];
      point[0][1] = params[1];
      point[1][0] = fabs(params[2]);
      point[1][1] = params[3];
      if (point[1][1] <= 0) return -4;
      sg = params[2];
    } else {
      point[0][0] = params[0];
      point[0][1] = params[1];
      point[0][2] = params[2];
      point[1][0] = fabs(params[3]);
      point[1][1] = params[4];
      if (point[1][1] <= 0) return -4;
      point[1][2] = params[5];
      if (point[1][2] <= 0) return -4;
      sg = params[3];
    }
    front[0] = (double)(sg > 0 ? 1 : sg < 0 ? -1 : 0);
  } else if (ps == OPS_CYL) {
    double lengthinv, len = 0;
    if (dim != 3) return -2;
    for (pt = 0; pt < 2; pt++)
      for (d = 0; d < 3; d++) point[pt][d] = params[3 * pt + d];
    if (point[0][0] == point[1][0] && point[0][1] == point[1][1] && point[0][2] == point[1][2]) return -5;
    point[2][0] = fabs(params[6]);
    point[2][1] = params[7];
    if (point[2][1] <= 0) return -4;
    point[2][2] = params[8];
    if (point[2][2] <= 0) return -4;
    front[2] = (double)(params[6] > 0 ? 1 : params[6] < 0 ? -1 : 0);
    for (d = 0; d < 3; d++) len += (point[1][d] - point[0][d]) * (point[1][d] - point[0][d]); /* Geo_LineLength */
    lengthinv = 1.0 / sqrt(len);
    for (d = 0; d < 3; d++) {
      point[3][d] = -(point[1][d] - point[0][d]) * lengthinv;
      point[4][d] = (point[1][d] - point[0][d]) * lengthinv;
    }
  } else if (ps == OPS_HEMI) {
    double len = 0;
    if (dim != 3) return -2;
    for (d = 0; d < 3; d++) point[0][d] = params[d];
    point[1][0] = fabs(params[3]);
    point[1][1] = params[7];
    if (point[1][1] <= 0) return -4;
    point[1][2] = params[8];
    if (point[1][2] <= 0) return -4;
    for (d = 0; d < 3; d++) point[2][d] = params[4 + d];
    for (d = 0; d < 3; d++) len += point[2][d] * point[2][d]; /* normalizeVD (Rn.c:500-508) */
    if (!len) return -6;
    len = sqrt(len);
    for (d = 0; d < 3; d++) point[2][d] /= len;
    front[0] = (double)(params[3] > 0 ? 1 : params[3] < 0 ? -1 : 0);
  } else if (ps == OPS_DISK) {
    double len;
    if (dim != 3) return -2;
    point[0][0] = params[0];
    point[0][1] = params[1];
    point[0][2] = params[2];
    point[1][0] = params[3];
    if (point[1][0] <= 0) return -7;
    point[1][1] = params[7];
    if (point[1][1] <= 0) return -4;
    front[0] = params[4];
    front[1] = params[5];
    front[2] = params[6];
    len = sqrt(front[0] * front[0] + front[1] * front[1] + front[2] * front[2]); /* normalizeVD */
    if (len <= 0) return -8;
    for (d = 0; d < 3; d++) front[d] /= len;
  } else return -2;
  if (srf->npanel[ps] % 64 == 0)
    srf->panels[ps] = (int*)realloc(srf->panels[ps], (srf->npanel[ps] + 64) * sizeof(int));
  srf->panels[ps][srf->npanel[ps]++] = ss->npnl;
  sim->ok = 0;
  return ss->npnl++;
}

/* jump statement: panel p1's face f1 jumps to panel p2's face f2 (and back
   when bidirectional), smolsurface.c surfsetjumppanel */
int osim_set_panel_jump(osim* sim, int p1, int f1, int p2, int f2, int bidirectional) {
  osurfss* ss = sim->srfss;
  if (!ss || p1 < 0 || p1 >= ss->npnl || p2 < 0 || p2 >= ss->npnl) return 1;
  if (!(f1 == OPF_FRONT || f1 == OPF_BACK) || !(f2 == OPF_FRONT || f2 == OPF_BACK)) return 2;
  ss->pnl[p1].jumpp[f1] = p2;
  ss->pnl[p1].jumpf[f1] = f2;
  if (bidirectional) {
    ss->pnl[p2].jumpp[f2] = p1;
    ss->pnl[p2].jumpf[f2] = f1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* panel functions                                                           */

/* panelside (smolsurface.c:3930-3963), strict 0 or 1, current positions */
static int os_panelside(const osim* sim, const double* pt, const opanel* pnl, double* distptr, int strict) {
  int dim = sim->dim, d, face;
  double dist = 0;
  if (pnl->ps == OPS_RECT) {
    d = (int)pnl->front[1];
    dist = pnl->front[0] * (pt[d] - pnl->point[0][d]);
  } else if (pnl->ps == OPS_TRI || pnl->ps == OPS_DISK) {
    for (d = 0; d < dim; d++) dist += (pt[d] - pnl->point[0][d]) * pnl->front[d];
  } else if (pnl->ps == OPS_SPH || pnl->ps == OPS_HEMI) {
    for (d = 0; d < dim; d++) dist += (pt[d] - pnl->point[0][d]) * (pt[d] - pnl->point[0][d]);
    dist = pnl->front[0] * (sqrt(dist) - pnl->point[1][0]);
  } else if (pnl->ps == OPS_CYL) {
    double cylnorm[3];
    dist = osurf_geo_linenormal3d(pnl->point[0], pnl->point[1], pt, cylnorm);
    dist = pnl->front[2] * (dist - pnl->point[2][0]);
  }
  if (dist > 0) face = OPF_FRONT;
  else if (!strict || dist < 0) face = OPF_BACK;
  else face = OPF_NONE;
  if (distptr) *distptr = dist;
  return face;
}

/* lineXpanel (smolsurface.c:3999-4101) */
static int os_lineXpanel(const osim* sim, const double* pt1, const double* pt2, const opanel* pnl, double* crsspt,
                         int* face1ptr, int* face2ptr, double* crossptr, double* cross2ptr) {
  int dim = sim->dim, d, intsct, face1, face2, facein;
  double dist1, dist2, cross = -1, cross2 = -1, nrdist, nrpos, dot;
  const double(*point)[3] = (const double(*)[3])pnl->point;
  const double* front = pnl->front;
  face1 = os_panelside(sim, pt1, pnl, &dist1, 0);
  face2 = os_panelside(sim, pt2, pnl, &dist2, 0);
  if (pnl->ps == OPS_RECT) {
    if (face1 == face2) return 0;
    cross = dist1 / (dist1 - dist2);
    for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    if (dim == 1) intsct = 1;
    else if (dim == 2) {
      d = (int)front[2];
      intsct = ((point[0][d] <= crsspt[d] && crsspt[d] <= point[1][d]) ||
                (point[1][d] <= crsspt[d] && crsspt[d] <= point[0][d]));
    } else {
      d = (int)front[2];
      intsct = ((point[0][d] <= crsspt[d] && crsspt[d] <= point[1][d]) ||
                (point[1][d] <= crsspt[d] && crsspt[d] <= point[0][d]));
      d = (d + 1) % 3;
      if (d == (int)front[1]) d = (d + 1) % 3;
      intsct = intsct && ((point[1][d] <= crsspt[d] && crsspt[d] <= point[2][d]) ||
                          (point[2][d] <= crsspt[d] && crsspt[d] <= point[1][d]));
    }
  } else if (pnl->ps == OPS_TRI) {
    if (face1 == face2) return 0;
    cross = dist1 / (dist1 - dist2);
    for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    if (dim == 1) intsct = 1;
    else if (dim == 3) intsct = osurf_geo_ptintriangle2(point, crsspt);
    else {
      double v1[2] = {point[1][0] - point[0][0], point[1][1] - point[0][1]};
      double v2[2] = {crsspt[0] - point[0][0], crsspt[1] - point[0][1]};
      double v3[2] = {crsspt[0] - point[1][0], crsspt[1] - point[1][1]};
      intsct = v1[0] * v2[0] + v1[1] * v2[1] >= 0;
      intsct = intsct && (v1[0] * v3[0] + v1[1] * v3[1] <= 0);
    }
  } else if (pnl->ps == OPS_SPH || pnl->ps == OPS_HEMI) {
    facein = front[0] > 0 ? OPF_BACK : OPF_FRONT;
    if (face1 == facein && face2 == facein) return 0;
    cross = osurf_geo_linexsphs(pt1, pt2, point[0], point[1][0], dim, &cross2, &nrdist, &nrpos);
    if (face1 == face2 && (!(nrdist <= point[1][0]) || !(nrpos >= 0) || !(nrpos <= 1))) return 0;
    if (face1 == facein) cross = cross2;
    for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    if (pnl->ps == OPS_SPH) intsct = 1;
    else { /* hemisphere: the crossing must be on the dome side of the opening */
      dot = 0;
      for (d = 0; d < dim; d++) dot += (crsspt[d] - point[0][d]) * point[2][d];
      intsct = (dot <= 0);
      if (!intsct && face1 == face2) {
        cross = cross2;
        face1 = (face1 == OPF_FRONT) ? OPF_BACK : OPF_FRONT;
        for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
        dot = 0;
        for (d = 0; d < dim; d++) dot += (crsspt[d] - point[0][d]) * point[2][d];
        intsct = (dot <= 0);
      }
    }
  } else if (pnl->ps == OPS_CYL) {
    facein = (int)front[2] > 0 ? OPF_BACK : OPF_FRONT;
    if (face1 == facein && face1 == face2) return 0;
    cross = osurf_geo_linexcyls(pt1, pt2, point[0], point[1], point[2][0], &cross2, &nrdist, &nrpos);
    if (face1 == face2 && (!(nrdist <= point[2][0]) || !(nrpos >= 0) || !(nrpos <= 1))) return 0;
    if (face1 == facein) cross = cross2;
    for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    intsct = osurf_geo_ptinslab(point[0], point[1], crsspt, dim);
    if (!intsct && face1 == face2) {
      cross = cross2;
      face1 = (face1 == OPF_FRONT) ? OPF_BACK : OPF_FRONT;
      for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
      intsct = osurf_geo_ptinslab(point[0], point[1], crsspt, dim);
    }
  } else if (pnl->ps == OPS_DISK) {
    if (face1 == face2) return 0;
    cross = dist1 / (dist1 - dist2);
    for (d = 0; d < dim; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    dot = 0;
    for (d = 0; d < dim; d++) dot += (crsspt[d] - point[0][d]) * (crsspt[d] - point[0][d]);
    intsct = (dot <= point[1][0] * point[1][0]);
  } else intsct = 0;
  if (face1ptr) *face1ptr = face1;
  if (face2ptr) *face2ptr = face2;
  if (crossptr) *crossptr = cross;
  if (cross2ptr) *cross2ptr = cross2;
  return intsct;
}

/* getpanelnormal (smolsurface.c:4430-4456) */
static void os_getpanelnormal(const osim* sim, const double* pt, const opanel* pnl, double* norm) {
  int dim = sim->dim, d;
  if (pnl->ps == OPS_RECT) {
    for (d = 0; d < dim; d++) norm[d] = 0;
    norm[(int)pnl->front[1]] = pnl->front[0];
  } else if (pnl->ps == OPS_TRI || pnl->ps == OPS_DISK) {
    for (d = 0; d < dim; d++) norm[d] = pnl->front[d];
  } else if (pnl->ps == OPS_SPH || pnl->ps == OPS_HEMI) {
    osurf_geo_spherenormal(pnl->point[0], pt, (int)pnl->front[0], dim, norm);
  } else if (pnl->ps == OPS_CYL) {
    osurf_geo_linenormal3d(pnl->point[0], pnl->point[1], pt, norm);
    if (pnl->front[2] == -1)
      for (d = 0; d < dim; d++) norm[d] *= -1;
  } else {
    for (d = 0; d < dim; d++) norm[d] = 0;
    norm[0] = 1;
  }
}

/* panelnormal (smolsurface.c:3967-3995) */
static void os_panelnormal(const osim* sim, const opanel* pnl, const double* pos, int face, double* norm) {
  int dim = sim->dim, d;
  if (face != OPF_BACK) face = OPF_FRONT;
  if (pnl->ps == OPS_RECT) {
    for (d = 0; d < dim; d++) norm[d] = 0;
    norm[(int)pnl->front[1]] =
        ((face == OPF_FRONT && pnl->front[0] == 1) || (face == OPF_BACK && pnl->front[0] == -1)) ? 1.0 : -1.0;
  } else if (pnl->ps == OPS_TRI || pnl->ps == OPS_DISK) {
    if (face == OPF_FRONT)
      for (d = 0; d < dim; d++) norm[d] = pnl->front[d];
    else
      for (d = 0; d < dim; d++) norm[d] = -pnl->front[d];
  } else if (pnl->ps == OPS_SPH || pnl->ps == OPS_HEMI) {
    osurf_geo_spherenormal(
        pnl->point[0], pos,
        ((face == OPF_FRONT && pnl->front[0] == 1) || (face == OPF_BACK && pnl->front[0] == -1)) ? 1 : -1, dim,
        norm);
  } else if (pnl->ps == OPS_CYL) {
    osurf_geo_linenormal3d(pnl->point[0], pnl->point[1], pos, norm);
    if ((face == OPF_BACK && pnl->front[2] == 1) || (face == OPF_FRONT && pnl->front[2] == -1))
      for (d = 0; d < dim; d++) norm[d] *= -1;
  }
}

/* fixpt2panel (smolsurface.c:4460-4474) */
static void os_fixpt2panel(const osim* sim, double* pt, const opanel* pnl, int face, double epsilon) {
  int dim = sim->dim, d, sign, faceat;
  double norm[3], dist, factor;
  faceat = os_panelside(sim, pt, pnl, &dist, 1);
  if ((faceat == face || face == OPF_NONE) && fabs(dist) <= epsilon) return;
  os_getpanelnormal(sim, pt, pnl, norm);
  for (d = 0; d < dim; d++) pt[d] -= dist * norm[d];
  if (face == OPF_NONE || face == OPF_BOTH) return;
  sign = (face == OPF_FRONT) ? 1 : -1;
  for (factor = 1.0; face != os_panelside(sim, pt, pnl, NULL, 1); factor *= 2.0)
    for (d = 0; d < dim; d++) pt[d] += sign * factor * DBL_EPSILON * norm[d];
}

/* ptinpanel (smolsurface.c:4276-4323) */
static int os_ptinpanel(const osim* sim, const double* pt, const opanel* pnl) {
  int dim = sim->dim, d;
  const double(*point)[3] = (const double(*)[3])pnl->point;
  if (pnl->ps == OPS_RECT) {
    int in;
    if (dim == 1) return 1;
    d = (int)pnl->front[2];
    in = ((point[0][d] <= pt[d] && pt[d] <= point[1][d]) || (point[1][d] <= pt[d] && pt[d] <= point[0][d]));
    if (dim == 2) return in;
    d = (d + 1) % 3;
    if (d == (int)pnl->front[1]) d = (d + 1) % 3;
    return in && ((point[1][d] <= pt[d] && pt[d] <= point[2][d]) || (point[2][d] <= pt[d] && pt[d] <= point[1][d]));
  } else if (pnl->ps == OPS_TRI) {
    if (dim == 1) return 1;
    if (dim == 3) return osurf_geo_ptintriangle2(point, pt);
    return osurf_geo_ptinslab(point[0], point[1], pt, 2);
  } else if (pnl->ps == OPS_SPH) {
    return 1;
  } else if (pnl->ps == OPS_HEMI) {
    double dot = 0;
    for (d = 0; d < dim; d++) dot += (pt[d] - point[0][d]) * point[2][d];
    return dot <= 0;
  } else if (pnl->ps == OPS_CYL) {
    return osurf_geo_ptinslab(point[0], point[1], pt, dim);
  } else if (pnl->ps == OPS_DISK) {
    double len2 = 0, dot = 0;
    for (d = 0; d < dim; d++) len2 += (pt[d] - point[0][d]) * (pt[d] - point[0][d]);
    if (len2 <= point[1][0] * point[1][0]) return 1;
    for (d = 0; d < dim; d++) dot += (pt[d] - point[0][d]) * pnl->front[d];
    len2 -= dot * dot;
    return len2 <= point[1][0] * point[1][0];
  }
  return 0;
}

/* surfacereflect (smolsurface.c:4821-4869), solution-phase molecules */
static void os_surfacereflect(const osim* sim, omol* mp, const opanel* pnl, const double* crsspt) {
  int dim = sim->dim, d, axis;
  double* pos = mp->pos;
  double norm[3], dot;
  if (pnl->ps == OPS_RECT) {
    axis = (int)pnl->front[1];
    pos[axis] -= 2.0 * (pos[axis] - crsspt[axis]);
  } else if (pnl->ps == OPS_TRI || pnl->ps == OPS_DISK) {
    dot = 0;
    for (d = 0; d < dim; d++) dot += (pos[d] - crsspt[d]) * pnl->front[d];
    for (d = 0; d < dim; d++) pos[d] -= 2.0 * pnl->front[d] * dot;
  } else if (pnl->ps == OPS_SPH || pnl->ps == OPS_HEMI || pnl->ps == OPS_CYL) {
    if (pnl->ps == OPS_CYL) osurf_geo_linenormal3d(pnl->point[0], pnl->point[1], crsspt, norm);
    else osurf_geo_spherenormal(pnl->point[0], crsspt, 1, dim, norm);
    dot = 0;
    for (d = 0; d < dim; d++) dot += (pos[d] - crsspt[d]) * norm[d];
    for (d = 0; d < dim; d++) pos[d] -= 2.0 * norm[d] * dot;
  }
}

/* surfacejump (smolsurface.c:4873-4926) */
static int os_surfacejump(osim* sim, omol* mp, const opanel* pnl, double* crsspt, int face) {
  osurfss* ss = sim->srfss;
  int dim = sim->dim, d, dir, p2 = pnl->jumpp[face], face2 = pnl->jumpf[face];
  const opanel* pnl2;
  double delta[3], cent[3], dot;
  if (p2 < 0) return 0;
  if (!(face2 == OPF_FRONT || face2 == OPF_BACK)) return 0;
  pnl2 = &ss->pnl[p2];
  if (pnl->ps == OPS_RECT) {
    /* Geo_RectCenter: mean of the corners (point[0] in 1D; the 2 ends in 2D;
       corners 0 and 2 in 3D) */
    if (dim == 1) {
      cent[0] = pnl->point[0][0];
      delta[0] = pnl2->point[0][0];
    } else if (dim == 2) {
      for (d = 0; d < 2; d++) {
        cent[d] = 0.5 * (pnl->point[0][d] + pnl->point[1][d]);
        delta[d] = 0.5 * (pnl2->point[0][d] + pnl2->point[1][d]);
      }
    } else {
      for (d = 0; d < 3; d++) {
        cent[d] = 0.5 * (pnl->point[0][d] + pnl->point[2][d]);
        delta[d] = 0.5 * (pnl2->point[0][d] + pnl2->point[2][d]);
      }
    }
    for (d = 0; d < dim; d++) delta[d] -= cent[d];
    dir = (pnl->front[0] == pnl2->front[0]) ? 1 : -1;
  } else if (pnl->ps == OPS_TRI) {
    /* Geo_TriCenter (Geometry.c:43-53): the vertex in 1D, 0.5 times the sum
       of the ends in 2D, (1.0/3.0) times the sum of the vertices in 3D */
    const double w = dim == 1 ? 1.0 : dim == 2 ? 0.5 : 1.0 / 3.0;
    for (d = 0; d < dim; d++) {
      int k;
      cent[d] = delta[d] = 0;
      for (k = 0; k < dim; k++) {
        cent[d] += pnl->point[k][d];
        delta[d] += pnl2->point[k][d];
      }
      cent[d] *= w;
      delta[d] *= w;
    }
    for (d = 0; d < dim; d++) delta[d] -= cent[d];
    dot = 0;
    for (d = 0; d < dim; d++) dot += pnl->front[d] * pnl2->front[d];
    dir = dot > 0 ? 1 : -1;
  } else if (pnl->ps == OPS_SPH || pnl->ps == OPS_HEMI) {
    for (d = 0; d < dim; d++)
      delta[d] = (crsspt[d] - pnl->point[0][d]) * pnl2->point[1][0] / pnl->point[1][0] + pnl2->point[0][d] - crsspt[d];
    dir = (pnl->front[0] == pnl2->front[0]) ? 1 : -1;
  } else if (pnl->ps == OPS_CYL) {
    for (d = 0; d < dim; d++)
      delta[d] = (crsspt[d] - pnl->point[0][d]) * pnl2->point[2][0] / pnl->point[2][0] + pnl2->point[0][d] - crsspt[d];
    dir = (pnl->front[2] == pnl2->front[2]) ? 1 : -1;
  } else if (pnl->ps == OPS_DISK) {
    for (d = 0; d < dim; d++) delta[d] = pnl2->point[0][d] - pnl->point[0][d];
    dot = 0;
    for (d = 0; d < dim; d++) dot += pnl->front[d] * pnl2->front[d];
    dir = dot > 0 ? 1 : -1;
  } else {
    for (d = 0; d < dim; d++) delta[d] = 0;
    dir = 1;
  }
  for (d = 0; d < dim; d++) {
    crsspt[d] += delta[d];
    mp->pos[d] += delta[d];
    mp->posoffset[d] -= delta[d];
  }
  os_fixpt2panel(sim, crsspt, pnl2, face2, 0);
  dir *= (face != face2) ? 1 : -1;
  if (dir == -1) os_surfacereflect(sim, mp, pnl2, crsspt);
  return 1;
}

/* ------------------------------------------------------------------------ */
/* actions                                                                   */

/* surfaction (smolsurface.c:4330-4384); draws one uniform when the action is
   rate based (sub is the counter-mode draw index) */
static int os_surfaction(osim* sim, const osurface* srf, int face, int ident, int ms, int* i2ptr, int* ms2ptr,
                         const omol* mp, uint32_t sub, double (*draw)(osim*, const omol*, uint32_t)) {
  int act, ms2, ms3, i2 = ident;
  const oactdet* det;
  ms2 = ms;
  act = srf->action[ident][ms][face];
  if (act == OSA_MULT) {
    double r;
    det = srf->det[ident][ms][face];
    r = draw(sim, mp, sub);
    ms2 = OMS_NONE;
    for (ms3 = 0; ms3 < O_MSMAX1 && ms2 == OMS_NONE; ms3++)
      if (r < det->cumprob[ms3]) ms2 = ms3;
    i2 = det->newspec[ms2];
    if (i2 == 0) act = OSA_ABSORB;
    else if (ms == OMS_SOLN) {
      if (face == OPF_FRONT) {
        if (ms2 == OMS_SOLN) act = OSA_REFLECT;
        else if (ms2 == OMS_BSOLN) act = OSA_TRANS;
        else act = OSA_ADSORB;
      } else {
        if (ms2 == OMS_SOLN) act = OSA_TRANS;
        else if (ms2 == OMS_BSOLN) act = OSA_REFLECT;
        else act = OSA_ADSORB;
      }
    } else {
      if (face == OPF_NONE) {
        if (ms2 == ms) act = OSA_NO;
        else if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) act = det->revprob[ms2] > 0 ? OSA_REVDES : OSA_IRREVDES;
        else act = OSA_FLIP;
      } else if (face == OPF_FRONT) {
        if (ms2 == OMS_SOLN) act = OSA_REFLECT;
        else if (ms2 == OMS_BSOLN) act = OSA_TRANS;
        else act = OSA_ADSORB;
      } else {
        if (ms2 == OMS_SOLN) act = OSA_TRANS;
        else if (ms2 == OMS_BSOLN) act = OSA_REFLECT;
        else act = OSA_ADSORB;
      }
    }
  } else {
    ms2 = (ms == OMS_BSOLN) ? OMS_SOLN : ms;
    det = srf->det[ident][ms2][face];
    if (det && det->datasrc[ms2] == 3) i2 = det->newspec[ms2];
  }
  if (i2ptr) *i2ptr = i2;
  if (ms2ptr) *ms2ptr = ms2;
  return act;
}

/* molchangeident (smolmolec.c:2863-2909) */
static void os_molchangeident(osim* sim, omol* mp, int ll, int mi, int i, int ms, int pnl, const double* crsspt) {
  osurfss* ss = sim->srfss;
  int dim = sim->dim, ll2, oldi, oldms, d;
  double epsilon = ss ? ss->epsilon : 0;
  if (i == 0) {
    osim_molkill(sim, mp, ll, mi);
    return;
  }
  oldi = mp->ident;
  oldms = mp->mstate;
  mp->ident = i;
  mp->mstate = ms;
  if (ms == OMS_SOLN || ms == OMS_BSOLN) mp->pnl = -1;
  else mp->pnl = pnl;
  if (crsspt && sim->difc[oldi][oldms] > 0) { /* correct motion for the new diffusion coefficient */
    double difcratio = sqrt(sim->difc[i][ms] / sim->difc[oldi][oldms]);
    for (d = 0; d < dim; d++) mp->pos[d] = crsspt[d] + difcratio * (mp->pos[d] - crsspt[d]);
  }
  if (ms == OMS_SOLN && mp->pnlx < 0) {
  } else if (ms == OMS_SOLN) {
    os_fixpt2panel(sim, mp->posx, &ss->pnl[mp->pnlx], OPF_FRONT, epsilon);
  } else if (ms == OMS_BSOLN && mp->pnlx < 0) {
    mp->mstate = OMS_SOLN;
  } else if (ms == OMS_BSOLN) {
    mp->mstate = OMS_SOLN;
    os_fixpt2panel(sim, mp->posx, &ss->pnl[mp->pnlx], OPF_BACK, epsilon);
  } else if (ms == OMS_FRONT) {
    os_fixpt2panel(sim, mp->pos, &ss->pnl[pnl], OPF_FRONT, epsilon);
  } else if (ms == OMS_BACK) {
    os_fixpt2panel(sim, mp->pos, &ss->pnl[pnl], OPF_BACK, epsilon);
  } else {
    os_fixpt2panel(sim, mp->pos, &ss->pnl[pnl], OPF_NONE, epsilon);
  }
  ll2 = sim->listlookup[i][ms];
  if (ll >= 0 && ll2 != ll) {
    mp->list = ll2;
    if (mi < 0) sim->sortl[ll] = 0;
    else if (mi < sim->sortl[ll]) sim->sortl[ll] = mi;
  }
}

void osim_molchangeident(osim* sim, omol* mp, int ll, int mi, int i, int ms, int pnl) {
  os_molchangeident(sim, mp, ll, mi, i, ms, pnl, NULL);
}

static int os_checksurfaces1mol(osim* sim, omol* mp, double crossminimum, int pass);

/* dosurfinteract (smolsurface.c:4930-5051).  Returns done. */
static int os_dosurfinteract(osim* sim, omol* mp, int ll, int mi, int p, int face, double* crsspt, uint32_t sub) {
  osurfss* ss = sim->srfss;
  const opanel* pnl = &ss->pnl[p];
  const osurface* srf = ss->srf[pnl->srf];
  int dim = sim->dim, d, done = 0, i = mp->ident, ms = mp->mstate, i2, ms2, act;
  double epsilon = ss->epsilon;
  act = os_surfaction(sim, srf, face, i, ms, &i2, &ms2, mp, sub, os_randCOD);
  if (act == OSA_NO) {
  } else if (act == OSA_TRANS) {
    int newface = (face == OPF_FRONT) ? OPF_BACK : OPF_FRONT;
    os_fixpt2panel(sim, crsspt, pnl, newface, epsilon);
    if (i2 != i) os_molchangeident(sim, mp, ll, mi, i2, ms, mp->pnl, crsspt);
  } else if (act == OSA_REFLECT) {
    if (ms != OMS_SOLN) return -2; /* surface-bound reflection needs movemol2closepanel */
    os_surfacereflect(sim, mp, pnl, crsspt);
    os_fixpt2panel(sim, crsspt, pnl, face, epsilon);
    if (i2 != i) os_molchangeident(sim, mp, ll, mi, i2, ms, mp->pnl, crsspt);
  } else if (act == OSA_ABSORB) {
    osim_molkill(sim, mp, ll, mi);
    done = 1;
  } else if (act == OSA_JUMP) {
    if (!os_surfacejump(sim, mp, pnl, crsspt, face)) {
      int newface = (face == OPF_FRONT) ? OPF_BACK : OPF_FRONT;
      os_fixpt2panel(sim, crsspt, pnl, newface, epsilon);
    }
    if (i2 != i) os_molchangeident(sim, mp, ll, mi, i2, ms, mp->pnl, crsspt);
  } else if (act == OSA_ADSORB) {
    int oldpnl = mp->pnl;
    os_molchangeident(sim, mp, ll, mi, i2, ms2, p, NULL);
    for (d = 0; d < dim; d++) mp->pos[d] = crsspt[d];
    if (!os_ptinpanel(sim, mp->pos, &ss->pnl[mp->pnl])) return -3; /* movept2panel not restated */
    if (ms == OMS_FRONT) os_fixpt2panel(sim, mp->pos, &ss->pnl[oldpnl], OPF_FRONT, epsilon);
    else if (ms == OMS_BACK) os_fixpt2panel(sim, mp->pos, &ss->pnl[oldpnl], OPF_BACK, epsilon);
    else if (ms == OMS_UP || ms == OMS_DOWN) return -3;
    if (ms2 == OMS_FRONT) os_fixpt2panel(sim, mp->pos, &ss->pnl[mp->pnl], OPF_FRONT, epsilon);
    else if (ms2 == OMS_BACK) os_fixpt2panel(sim, mp->pos, &ss->pnl[mp->pnl], OPF_BACK, epsilon);
    done = 1;
  } else if (act == OSA_REVDES || act == OSA_IRREVDES) {
    omol* m2;
    double x, norm[3], y;
    int dface = (ms2 == OMS_SOLN) ? OPF_FRONT : OPF_BACK;
    for (d = 0; d < dim; d++) crsspt[d] = mp->pos[d];
    os_fixpt2panel(sim, crsspt, pnl, dface, epsilon);
    m2 = osim_getnextmol(sim);
    if (!m2) return -1;
    m2->ident = i2;
    m2->mstate = OMS_SOLN;
    m2->serno = mp->serno;
    y = os_randCCD(sim, mp, (SMOLRNG_SURF_DESORBDIST << 8));
    x = osurf_desorbdist_y(sim->difstep[i2][OMS_SOLN], y, act == OSA_REVDES);
    os_panelnormal(sim, pnl, crsspt, dface, norm);
    for (d = 0; d < dim; d++) {
      m2->posx[d] = m2->via[d] = crsspt[d];
      m2->pos[d] = crsspt[d] + x * norm[d];
      m2->posoffset[d] = mp->posoffset[d];
    }
    m2->pnl = m2->pnlx = -1;
    m2->box = osim_pos2box(sim, m2->pos);
    m2->list = sim->listlookup[i2][OMS_SOLN];
    osim_molkill(sim, mp, ll, mi);
    if (os_checksurfaces1mol(sim, m2, 0, SMOLRNG_SURF_DESORBED) < 0) return -2;
    done = 1;
  } else if (act == OSA_FLIP) { /* on-surface flip (smolsurface.c:5044-5048) */
    os_molchangeident(sim, mp, ll, mi, i2, ms2, p, crsspt);
    if (ms2 == OMS_FRONT) os_fixpt2panel(sim, mp->pos, &ss->pnl[mp->pnl], OPF_FRONT, epsilon);
    else if (ms2 == OMS_BACK) os_fixpt2panel(sim, mp->pos, &ss->pnl[mp->pnl], OPF_BACK, epsilon);
    done = 1;
  } else {
    return -4; /* port, emitters: not restated */
  }
  return done;
}

/* line2nextbox (smolboxes.c:846-892) */
static int os_line2nextbox(const osim* sim, const double* pt1, const double* pt2, int b) {
  int dim = sim->dim, d, d2 = 0, boxside, boxside2 = 0, sum, flag = 0, z1[O_DIMMAX], adrs;
  const obox* bp = &sim->boxes[b];
  double crsmin = 1.01, edge, crs;
  if (osim_pos2box(sim, pt2) == b) return -1;
  for (d = 0; d < dim; d++) {
    z1[d] = bp->indx[d];
    if (pt2[d] != pt1[d]) {
      boxside = (pt2[d] > pt1[d]) ? 1 : 0;
      sum = bp->indx[d] + boxside;
      if (sum > 0 && sum < sim->side[d]) {
        edge = sim->bmin[d] + (double)sum * sim->bsize[d];
        crs = (edge - pt1[d]) / (pt2[d] - pt1[d]);
        if (crs < crsmin) {
          crsmin = crs;
          d2 = d;
          boxside2 = boxside;
          flag = 0;
        } else if (crs == crsmin) {
          if (boxside == 0 && boxside2 == 0 && (flag == 0 || flag == 2)) flag = 2;
          else flag = 1;
        }
      }
    }
  }
  if (flag) {
    for (d = 0; d < dim; d++) {
      if (pt2[d] != pt1[d]) {
        boxside = (pt2[d] > pt1[d]) ? 1 : 0;
        sum = bp->indx[d] + boxside;
        if (sum > 0 && sum < sim->side[d]) {
          edge = sim->bmin[d] + (double)sum * sim->bsize[d];
          crs = (edge - pt1[d]) / (pt2[d] - pt1[d]);
          if (crs == crsmin && (boxside == 1 || flag == 2)) z1[d] += boxside ? 1 : -1;
        }
      }
    }
  } else {
    if (crsmin == 1.01) return -1;
    z1[d2] += boxside2 ? 1 : -1;
  }
  adrs = z1[0]; /* indx2addZV */
  for (d = 1; d < dim; d++) adrs = sim->side[d] * adrs + z1[d];
  return adrs;
}

/* the crossing search of checksurfaces / checksurfaces1mol: nearest panel
   crossed by via->pos over the boxes the segment passes */
static int os_nearest(osim* sim, omol* mp, double crossminimum, int use_min, double* crssptmin, int* facemin,
                      double* crossminp, double* crossmin2p) {
  osurfss* ss = sim->srfss;
  int dim = sim->dim, b, k, d, pnlmin = -1, face;
  double crossmin = 2, crossmin2 = 2, cross, crsspt[3];
  *facemin = OPF_FRONT;
  for (b = osim_pos2box(sim, mp->via); b >= 0; b = os_line2nextbox(sim, mp->via, mp->pos, b)) {
    for (k = ss->bpstart[b]; k < ss->bpstart[b + 1]; k++) {
      int p = ss->bplist[k];
      if (p == mp->pnl) continue;
      if (os_lineXpanel(sim, mp->via, mp->pos, &ss->pnl[p], crsspt, &face, NULL, &cross, NULL) &&
          cross <= crossmin2 && (!use_min || cross - crossminimum > OS_VERYCLOSE)) {
        if (cross <= crossmin) {
          crossmin2 = crossmin;
          crossmin = cross;
          pnlmin = p;
          for (d = 0; d < dim; d++) crssptmin[d] = crsspt[d];
          *facemin = face;
        } else crossmin2 = cross;
      }
    }
  }
  *crossminp = crossmin;
  *crossmin2p = crossmin2;
  return pnlmin;
}

/* checksurfaces1mol (smolsurface.c:5055-5112) */
static int os_checksurfaces1mol(osim* sim, omol* mp, double crossminimum, int pass) {
  int dim = sim->dim, d, done = 0, it = 0, facemin, p;
  double crossmin, crossmin2, crssptmin[3];
  while (!done) {
    if (++it > 100) {
      for (d = 0; d < dim; d++) mp->pos[d] = mp->posx[d];
      break;
    }
    p = os_nearest(sim, mp, crossminimum, 1, crssptmin, &facemin, &crossmin, &crossmin2);
    if (crossmin < 2) {
      int flag = (crossmin2 != crossmin && crossmin2 - crossmin < OS_VERYCLOSE) ? 1 : 0;
      if (flag) {
        for (d = 0; d < dim; d++) mp->pos[d] = mp->via[d];
        done = 1;
      } else {
        int ll = sim->listsystem[mp->list] ? mp->list : -1;
        done = os_dosurfinteract(sim, mp, ll, -1, p, facemin, crssptmin, ((uint32_t)pass << 8) | (uint32_t)it);
        if (done < 0) return done;
        for (d = 0; d < dim; d++) mp->via[d] = crssptmin[d];
      }
    } else done = 1;
  }
  return 0;
}

/* checksurfaces (smolsurface.c:5118-5180) */
int osurf_checksurfaces(osim* sim, int ll, int reborn) {
  osurfss* ss = sim->srfss;
  int dim = sim->dim, nmol, m, d;
  if (!ss || !ss->bpstart) return 0;
  if (!reborn && !(ss->srfmollist[ll] & OSML_DIFFUSE)) return 0;
  nmol = sim->nl[ll];
  m = reborn ? sim->topl[ll] : 0;
  for (; m < nmol; m++) {
    omol* mp = sim->live[ll][m];
    int done = 0, it = 0, facemin, p;
    double crossmin, crossmin2, crssptmin[3];
    for (d = 0; d < dim; d++) mp->via[d] = mp->posx[d];
    while (!done) {
      if (++it > 100) {
        for (d = 0; d < dim; d++) mp->pos[d] = mp->posx[d];
        break;
      }
      p = os_nearest(sim, mp, 0, 0, crssptmin, &facemin, &crossmin, &crossmin2);
      if (crossmin < 2) {
        int flag = (crossmin2 != crossmin && crossmin2 - crossmin < OS_VERYCLOSE) ? 1 : 0;
        if (flag) {
          for (d = 0; d < dim; d++) mp->pos[d] = mp->via[d];
          done = 1;
        } else {
          uint32_t sub = ((uint32_t)(reborn ? SMOLRNG_SURF_REBORN : SMOLRNG_SURF_DIFFUSED) << 8) | (uint32_t)it;
          done = os_dosurfinteract(sim, mp, ll, m, p, facemin, crssptmin, sub);
          if (done < 0) return done;
          for (d = 0; d < dim; d++) mp->via[d] = crssptmin[d];
        }
      } else done = 1;
    }
  }
  return 0;
}

/* checksurfacebound (smolsurface.c:5184-5199) */
int osurf_checksurfacebound(osim* sim, int ll) {
  osurfss* ss = sim->srfss;
  int m, nmol;
  if (!ss || !(ss->srfmollist[ll] & OSML_SRFBOUND)) return 0;
  nmol = sim->nl[ll];
  for (m = 0; m < nmol; m++) {
    omol* mp = sim->live[ll][m];
    if (mp->mstate != OMS_SOLN) {
      const opanel* pnl = &ss->pnl[mp->pnl];
      int face = OPF_NONE, act, i2, ms2;
      /* dosurfinteract with face PFnone: only rate-based actions apply */
      act = os_surfaction(sim, ss->srf[pnl->srf], face, mp->ident, mp->mstate, &i2, &ms2, mp,
                          (SMOLRNG_SURF_BOUND << 8), os_randCOD);
      if (act == OSA_NO) continue;
      if (act == OSA_REVDES || act == OSA_IRREVDES) {
        /* re-enter the desorption branch of dosurfinteract with the action
           already drawn */
        int d, dim = sim->dim;
        omol* m2;
        double x, norm[3], y, crsspt[3];
        int dface = (ms2 == OMS_SOLN) ? OPF_FRONT : OPF_BACK;
        for (d = 0; d < dim; d++) crsspt[d] = mp->pos[d];
        os_fixpt2panel(sim, crsspt, pnl, dface, ss->epsilon);
        m2 = osim_getnextmol(sim);
        if (!m2) return -1;
        m2->ident = i2;
        m2->mstate = OMS_SOLN;
        m2->serno = mp->serno;
        y = os_randCCD(sim, mp, (SMOLRNG_SURF_DESORBDIST << 8));
        x = osurf_desorbdist_y(sim->difstep[i2][OMS_SOLN], y, act == OSA_REVDES);
        os_panelnormal(sim, pnl, crsspt, dface, norm);
        for (d = 0; d < dim; d++) {
          m2->posx[d] = m2->via[d] = crsspt[d];
          m2->pos[d] = crsspt[d] + x * norm[d];
          m2->posoffset[d] = mp->posoffset[d];
        }
        m2->pnl = m2->pnlx = -1;
        m2->box = osim_pos2box(sim, m2->pos);
        m2->list = sim->listlookup[i2][OMS_SOLN];
        osim_molkill(sim, mp, ll, m);
        if (os_checksurfaces1mol(sim, m2, 0, SMOLRNG_SURF_DESORBED) < 0) return -2;
      } else if (act == OSA_ABSORB) {
        osim_molkill(sim, mp, ll, m);
      } else if (act == OSA_FLIP) { /* on-surface flip (smolsurface.c:5044-5048), crsspt = posx */
        os_molchangeident(sim, mp, ll, m, i2, ms2, mp->pnl, mp->posx);
        if (ms2 == OMS_FRONT) os_fixpt2panel(sim, mp->pos, &ss->pnl[mp->pnl], OPF_FRONT, ss->epsilon);
        else if (ms2 == OMS_BACK) os_fixpt2panel(sim, mp->pos, &ss->pnl[mp->pnl], OPF_BACK, ss->epsilon);
      } else {
        return -4;
      }
    }
  }
  return 0;
}

/* rxnXsurface (smolsurface.c:4388-4424): 1 when a surface between the two
   molecules blocks their reaction.  In counter mode the surface-action draws
   are keyed by the pair and a running index c. */
typedef struct { osim* sim; uint64_t s1, s2; uint32_t j; uint32_t c; } os_pairdraw;
static os_pairdraw* os_pd; /* set around each call */
static double os_randCOD_pair(osim* sim, const omol* mp, uint32_t sub) {
  (void)mp;
  (void)sub;
  if (sim->rngmode == ORNG_SFMT) return osim_randCOD(sim);
  {
    smolrng4 r = smolrng_pair(sim->key0, sim->key1, os_pd->s1, os_pd->s2, sim->stepindex, SMOLRNG_RXNSURF,
                              (os_pd->j << 8) | os_pd->c);
    os_pd->c++;
    return smolrng_u01(r.v[0]);
  }
}

static int os_trans(osim* sim, const osurface* srf, int face, int ident, int ms) {
  return os_surfaction(sim, srf, face, ident, ms, NULL, NULL, NULL, 0, os_randCOD_pair) == OSA_TRANS;
}

int osurf_rxnXsurface_j(osim* sim, omol* m1, omol* m2, int j) {
  osurfss* ss = sim->srfss;
  int b, k, result = 0, i1, i2, ms1, ms2, face1, face2, facein;
  double *pos1, *pos2, dc1, dc2, rxnpt, crsspt[3], cross, cross2;
  os_pairdraw pd;
  if (!ss || !ss->bpstart) return 0;
  pd.sim = sim;
  pd.s1 = m1->serno;
  pd.s2 = m2->serno;
  pd.j = (uint32_t)j;
  pd.c = 0;
  os_pd = &pd;
  i1 = m1->ident;
  i2 = m2->ident;
  ms1 = m1->mstate;
  ms2 = m2->mstate;
  pos1 = m1->pos;
  pos2 = m2->pos;
  dc1 = sim->difc[i1][ms1];
  dc2 = sim->difc[i2][ms2];
  if (dc1 == 0 && dc2 == 0) dc1 = dc2 = 1;
  rxnpt = dc1 / (dc1 + dc2);
  for (b = osim_pos2box(sim, m1->pos); b >= 0 && result == 0; b = os_line2nextbox(sim, pos1, pos2, b)) {
    for (k = ss->bpstart[b]; k < ss->bpstart[b + 1] && result == 0; k++) {
      int p = ss->bplist[k];
      const opanel* pnl = &ss->pnl[p];
      const osurface* srf = ss->srf[pnl->srf];
      if (os_lineXpanel(sim, pos1, pos2, pnl, crsspt, &face1, &face2, &cross, &cross2)) {
        if ((ms1 == OMS_UP || ms1 == OMS_DOWN) && p == m1->pnl) result = 0;
        else if ((ms2 == OMS_UP || ms2 == OMS_DOWN) && p == m2->pnl) result = 0;
        else if (face1 != face2) {
          if (rxnpt < cross || (rxnpt == cross && face1 == OPF_BACK)) result = !os_trans(sim, srf, face2, i2, ms2);
          else result = !os_trans(sim, srf, face1, i1, ms1);
        } else {
          facein = (face1 == OPF_FRONT) ? OPF_BACK : OPF_FRONT;
          if (rxnpt < cross || (rxnpt == cross && face1 == OPF_BACK))
            result = !(os_trans(sim, srf, face2, i2, ms2) && os_trans(sim, srf, facein, i2, ms2));
          else if (rxnpt < cross2 || (rxnpt == cross2 && face1 == OPF_FRONT))
            result = !(os_trans(sim, srf, face1, i1, ms1) && os_trans(sim, srf, face2, i2, ms2));
          else result = !(os_trans(sim, srf, face1, i1, ms1) && os_trans(sim, srf, facein, i1, ms1));
        }
      }
    }
  }
  os_pd = NULL;
  return result;
}

int osurf_rxnXsurface(osim* sim, omol* m1, omol* m2) { return osurf_rxnXsurface_j(sim, m1, m2, 0); }

/* ------------------------------------------------------------------------ */
/* update                                                                    */

/* srfsamestate (smolsurface.c:638-658) */
static int os_samestate(int ms1, int face1, int ms2, int* ms3p) {
  int same;
  if (face1 == OPF_FRONT && ms2 == OMS_SOLN) same = 1;
  else if (face1 == OPF_BACK && ms2 == OMS_BSOLN) same = 1;
  else if (face1 == OPF_NONE && ms1 == ms2) same = 1;
  else same = 0;
  if (ms3p) {
    if (face1 == OPF_FRONT) *ms3p = OMS_SOLN;
    else if (face1 == OPF_BACK) *ms3p = OMS_BSOLN;
    else *ms3p = (ms1 == OMS_SOLN) ? OMS_NONE : ms1;
  }
  return same;
}

/* srfreverseaction (smolsurface.c:661-710) */
static void os_reverseaction(int ms1, int face1, int ms2, int* ms3p, int* face2p, int* ms4p, int* actp) {
  int ms3, face2, ms4, act;
  if (ms1 == OMS_SOLN && face1 == OPF_NONE) {
    ms3 = ms4 = OMS_NONE;
    face2 = OPF_NONE;
    act = OSA_NONE;
  } else if (ms1 == OMS_SOLN) {
    if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) {
      ms3 = OMS_SOLN;
      face2 = (ms2 == OMS_SOLN) ? OPF_FRONT : OPF_BACK;
      act = (face1 == face2) ? OSA_REFLECT : OSA_TRANS;
    } else {
      ms3 = ms2;
      face2 = OPF_NONE;
      act = OSA_ADSORB;
    }
    ms4 = (face1 == OPF_FRONT) ? OMS_SOLN : OMS_BSOLN;
  } else if (face1 == OPF_NONE) {
    if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) {
      ms3 = OMS_SOLN;
      face2 = (ms2 == OMS_SOLN) ? OPF_FRONT : OPF_BACK;
      act = OSA_REVDES;
    } else {
      ms3 = ms2;
      face2 = OPF_NONE;
      act = OSA_FLIP;
    }
    ms4 = ms1;
  } else {
    if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) {
      ms3 = ms1;
      face2 = (ms2 == OMS_SOLN) ? OPF_FRONT : OPF_BACK;
      ms4 = (face1 == OPF_FRONT) ? OMS_SOLN : OMS_BSOLN;
      act = (face1 == face2) ? OSA_REFLECT : OSA_TRANS;
    } else {
      ms3 = ms2;
      face2 = OPF_BOTH;
      ms4 = ms1;
      act = OSA_NO;
    }
  }
  if (ms3p) *ms3p = ms3;
  if (face2p) *face2p = face2;
  if (ms4p) *ms4p = ms4;
  if (actp) *actp = act;
}

/* srfcalcprob (smolsurface.c:3070-3146).  Returns -3 where the reference
   leaves the probability undefined or finds it with random draws: a
   two-sided transmission given by rates with one side maximal and the other
   finite, or with a zero diffusion coefficient (surfacetransmit). */
static double os_calcprob(osim* sim, const osurface* srf, int i, int ms1, int face, int ms2) {
  const oactdet* a;
  double rate, raterev, sum, dt, difc, difcrev, prob;
  int ms3, face2, ms4, actrev, i2, ms5;
  if (ms1 == OMS_SOLN && face == OPF_NONE) return 0;
  if (srf->action[i][ms1][face] != OSA_MULT) return -1;
  a = srf->det[i][ms1][face];
  if (!a) return -1;
  if (a->datasrc[ms2] == 2) return a->prob[ms2];
  if (a->datasrc[ms2] != 1) return 0;
  rate = a->rate[ms2];
  if (rate < 0 && rate != -1) return -2;
  if (rate == 0) return 0;
  os_reverseaction(ms1, face, ms2, &ms3, &face2, &ms4, &actrev);
  i2 = a->newspec[ms2];
  if (!i2) raterev = 0;
  else if (actrev == OSA_NO) raterev = 0;
  else if (srf->action[i2][ms3][face2] == OSA_MULT) {
    const oactdet* b = srf->det[i2][ms3][face2];
    if (!b) raterev = 0;
    else if (b->datasrc[ms4] == 1) raterev = b->rate[ms4];
    else raterev = 0;
  } else if (srf->action[i2][ms3][face2] == actrev) raterev = -1;
  else raterev = 0;
  if (raterev < 0 && raterev != -1) raterev = 0;
  dt = sim->dt;
  difc = sim->difc[i][OMS_SOLN];
  difcrev = sim->difc[i2][OMS_SOLN];
  sum = 0;
  if (ms1 != OMS_SOLN && face == OPF_NONE)
    for (ms5 = 0; ms5 < O_MSMAX1; ms5++)
      if (ms5 != ms1 && a->datasrc[ms5] == 1 && a->rate[ms5] >= 0) sum += a->rate[ms5];
  if (ms1 == OMS_SOLN) {
    if ((face == OPF_FRONT && ms2 == OMS_SOLN) || (face == OPF_BACK && ms2 == OMS_BSOLN)) prob = 0;
    else if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) {
      if (raterev == 0) prob = osurf_surfaceprob(rate, 0, dt, difc, NULL, 0 /* SPAirrTrans */);
      else {
        prob = -2;
        if (osurf_transmit(rate, raterev, difc, difcrev, dt, &prob, NULL)) return -3;
      }
    } else {
      if (raterev == 0) prob = osurf_surfaceprob(rate, 0, dt, difc, NULL, 4 /* SPAirrAds */);
      else prob = osurf_surfaceprob(rate, raterev, dt, difc, NULL, 8 /* SPArevAds */);
    }
  } else if (face == OPF_NONE) {
    if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) {
      if (raterev == 0) prob = osurf_surfaceprob(rate, sum, dt, difc, NULL, 10 /* SPAirrDes */);
      else prob = osurf_surfaceprob(rate, raterev, dt, difc, NULL, 11 /* SPArevDes */);
    } else {
      if (ms1 == ms2) prob = 0;
      else if (raterev == 0) prob = osurf_surfaceprob(rate, sum, dt, difc, NULL, 12 /* SPAirrFlip */);
      else prob = osurf_surfaceprob(rate, raterev, dt, difc, NULL, 13 /* SPArevFlip */);
    }
  } else {
    if ((face == OPF_FRONT && ms2 == OMS_SOLN) || (face == OPF_BACK && ms2 == OMS_BSOLN)) prob = 0;
    else if (ms2 == OMS_SOLN || ms2 == OMS_BSOLN) {
      if (raterev == 0) prob = osurf_surfaceprob(rate, 0, dt, difc, NULL, 0 /* SPAirrTrans */);
      else {
        prob = -2;
        if (osurf_transmit(rate, raterev, difc, difcrev, dt, &prob, NULL)) return -3;
      }
    } else prob = osurf_surfaceprob(rate, 0, dt, difc, NULL, 4 /* SPAirrAds */);
  }
  return prob;
}

/* panelinbox (smolboxes.c:95-152) */
static int os_panelinbox(const osim* sim, const opanel* pnl, int b) {
  int dim = sim->dim, d;
  const obox* bp = &sim->boxes[b];
  double v1[3], v2[3];
  for (d = 0; d < dim; d++) { /* box2pos, smolboxes.c:54-63 */
    v1[d] = sim->bmin[d] + bp->indx[d] * sim->bsize[d];
    v2[d] = sim->bmin[d] + (bp->indx[d] + 1) * sim->bsize[d];
  }
  for (d = 0; d < dim; d++) {
    if (bp->indx[d] == 0) v1[d] = -OS_VERYLARGE;
    if (bp->indx[d] == sim->side[d] - 1) v2[d] = OS_VERYLARGE;
  }
  if (pnl->ps == OPS_RECT || pnl->ps == OPS_TRI) {
    if (dim == 1) return osurf_geo_ptinslab(v1, v2, pnl->point[0], dim);
    if (dim == 2) {
      double v3[2];
      if (pnl->ps == OPS_TRI) return osurf_geo_linexaabb2(pnl->point[0], pnl->point[1], pnl->front, v1, v2);
      v3[0] = pnl->front[1] == 0 ? 1 : 0;
      v3[1] = pnl->front[1] == 0 ? 0 : 1;
      return osurf_geo_linexaabb2(pnl->point[0], pnl->point[1], v3, v1, v2);
    }
    if (pnl->ps == OPS_RECT)
      return osurf_geo_rectxaabb3(pnl->point[0], pnl->point[1], pnl->point[3], pnl->point[0], v1, v2);
    return osurf_geo_trixaabb3(pnl->point[0], pnl->point[1], pnl->point[2], pnl->front, v1, v2);
  } else if (pnl->ps == OPS_SPH) {
    const double(*point)[3] = (const double(*)[3])pnl->point;
    if (dim == 1) {
      if ((point[0][0] - point[1][0] < v1[0] || point[0][0] - point[1][0] >= v2[0]) &&
          (point[0][0] + point[1][0] < v1[0] || point[0][0] + point[1][0] >= v2[0]))
        return 0;
      return 1;
    }
    if (dim == 2) return osurf_geo_circlexaabb2(point[0], point[1][0], v1, v2);
    return osurf_geo_sphsxaabb3(point[0], point[1][0], v1, v2);
  } else if (pnl->ps == OPS_CYL) {
    return osurf_geo_cylsxaabb3(pnl->point[0], pnl->point[1], pnl->point[2][0], v1, v2);
  } else if (pnl->ps == OPS_HEMI) {
    return osurf_geo_hemisxaabb3(pnl->point[0], pnl->point[1][0], pnl->point[2], v1, v2);
  } else if (pnl->ps == OPS_DISK) {
    return osurf_geo_diskxaabb3(pnl->point[0], pnl->point[1][0], pnl->front, v1, v2);
  }
  return -1;
}

/* surfupdatelists + surfupdateparams (smolsurface.c:3728-3860) and the box
   panel lists of boxesupdateparams (smolboxes.c:652-671) */
int osurf_update(osim* sim) {
  osurfss* ss = sim->srfss;
  int s, i, ms, f, ms2, ll, b, k, check = 0;
  if (!ss) return 0;
  /* srfmollist flags */
  for (s = 0; s < ss->nsrf && !check; s++)
    for (i = 1; i < sim->nspecies && !check; i++)
      for (ms = 0; ms < O_MSMAX && !check; ms++)
        for (f = 0; f <= OPF_NONE && !check; f++) {
          int act = ss->srf[s]->action[i][ms][f];
          if (!(act == OSA_NO || act == OSA_TRANS) || ss->srf[s]->det[i][ms][f]) check = 1;
        }
  for (i = 1; i < sim->nspecies; i++)
    for (ms = 0; ms < O_MSMAX; ms++) {
      ll = sim->listlookup[i][ms];
      if (ll < 0) continue;
      if (check && sim->difc[i][ms] > 0) ss->srfmollist[ll] |= OSML_DIFFUSE;
      if (ms != OMS_SOLN) ss->srfmollist[ll] |= OSML_SRFBOUND;
    }
  /* probabilities */
  for (s = 0; s < ss->nsrf; s++) {
    osurface* srf = ss->srf[s];
    for (i = 1; i < sim->nspecies; i++)
      for (ms = 0; ms < O_MSMAX; ms++)
        for (f = 0; f < 3; f++)
          if (srf->action[i][ms][f] == OSA_MULT && srf->det[i][ms][f]) {
            oactdet* a = srf->det[i][ms][f];
            for (ms2 = 0; ms2 < O_MSMAX1; ms2++) {
              double p = os_calcprob(sim, srf, i, ms, f, ms2);
              if (p == -3) return 3;
              a->prob[ms2] = p;
            }
          }
    for (i = 1; i < sim->nspecies; i++)
      for (ms = 0; ms < O_MSMAX; ms++)
        for (f = 0; f < 3; f++)
          if (srf->action[i][ms][f] == OSA_MULT && srf->det[i][ms][f]) {
            oactdet* a = srf->det[i][ms][f];
            double sum = 0;
            int ms3;
            for (ms2 = 0; ms2 < O_MSMAX1; ms2++)
              if (!os_samestate(ms, f, ms2, NULL)) sum += a->prob[ms2];
            if (sum > 1.0) {
              a->prob[ms] = 0;
              for (ms2 = 0; ms2 < O_MSMAX1; ms2++) a->prob[ms2] /= sum;
            } else {
              os_samestate(ms, f, OMS_SOLN, &ms3);
              a->prob[ms3] = 1.0 - sum;
            }
            sum = 0;
            for (ms2 = 0; ms2 < O_MSMAX1; ms2++) {
              sum += a->prob[ms2];
              a->cumprob[ms2] = sum;
            }
          }
    for (i = 1; i < sim->nspecies; i++)
      for (ms = 0; ms < O_MSMAX; ms++)
        for (f = 0; f < 3; f++)
          if (srf->action[i][ms][f] == OSA_MULT && srf->det[i][ms][f]) {
            oactdet* a = srf->det[i][ms][f];
            for (ms2 = 0; ms2 < O_MSMAX1; ms2++) {
              int ms3, face2, ms4, i2;
              os_reverseaction(ms, f, ms2, &ms3, &face2, &ms4, NULL);
              i2 = a->newspec[ms2];
              if (face2 != OPF_BOTH && ms3 < O_MSMAX && face2 < 3 && srf->det[i][ms3][face2] && srf->det[i2][ms3][face2])
                a->revprob[ms2] = srf->det[i2][ms3][face2]->prob[ms4];
            }
          }
  }
  /* box panel lists: surfaces in order, shapes in enum order, panels in
     creation order */
  free(ss->bpstart);
  free(ss->bplist);
  ss->bpstart = (int*)calloc(sim->nbox + 1, sizeof(int));
  {
    int cap = 64, n = 0, ps, q;
    ss->bplist = (int*)malloc(cap * sizeof(int));
    for (b = 0; b < sim->nbox; b++) {
      ss->bpstart[b] = n;
      for (s = 0; s < ss->nsrf; s++)
        for (ps = 0; ps < OPS_MAX; ps++)
          for (q = 0; q < ss->srf[s]->npanel[ps]; q++) {
            int p = ss->srf[s]->panels[ps][q];
            int in = os_panelinbox(sim, &ss->pnl[p], b);
            if (in < 0) return 4;
            if (in) {
              if (n == cap) {
                cap *= 2;
                ss->bplist = (int*)realloc(ss->bplist, cap * sizeof(int));
              }
              ss->bplist[n++] = p;
            }
          }
    }
    ss->bpstart[sim->nbox] = n;
  }
  (void)k;
  return 0;
}

/* molecules on surfaces: per-surface molecule lists are not kept (they only
   serve commands outside the particle step) */
void osurf_removemol(osim* sim, omol* mp, int ll) {
  (void)sim;
  (void)mp;
  (void)ll;
}
int osurf_addmol(osim* sim, omol* mp, int ll) {
  (void)sim;
  (void)mp;
  (void)ll;
  return 0;
}
void osurf_clearlist(osim* sim, int ll) {
  (void)sim;
  (void)ll;
}

/* movemol2closepanel (smolsurface.c:4726-4817) is needed only for mobile
   surface-bound molecules, which osurf_update refuses */
void osurf_movemol2closepanel(osim* sim, omol* mp) {
  (void)sim;
  (void)mp;
}

/* inspection: number of panels in box b (-1 without surfaces) */
int osim_box_panels(osim* sim, int b, int* out, int maxn) {
  osurfss* ss = sim->srfss;
  int k, n;
  if (!ss || !ss->bpstart || b < 0 || b >= sim->nbox) return -1;
  n = ss->bpstart[b + 1] - ss->bpstart[b];
  for (k = 0; k < n && k < maxn; k++) out[k] = ss->bplist[ss->bpstart[b] + k];
  return n;
}
