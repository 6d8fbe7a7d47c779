// Host side of the surfaces (see surfaces_host.h).  Panel geometry follows
// surfaddpanel (smolsurface.c:2432-2640) for 3D rectangles, spheres and
// disks; the box panel lists use the reference's separating-axis tests
// (Geometry.c Geo_RectXaabb3 :1640, Geo_SphsXaabb3 :1736, Geo_DiskXaabb3
// :1918) so each box lists exactly the panels the reference lists, in its
// order (surfaces, then shapes, then panels).
#include "surfaces_host.h"

#include "surfparam.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

namespace smolamd {

namespace {

enum { MSsoln_ = 0, MSbsoln_ = 5, MSnone_ = 7, MSall_ = 6 };
enum { PFfront_ = 0, PFback_ = 1, PFnone_ = 2, PFboth_ = 3 };
enum { SAreflect_ = 0, SAtrans_ = 1, SAabsorb_ = 2, SAjump_ = 3, SAport_ = 4, SAmult_ = 5, SAno_ = 6,
       SAnone_ = 7, SAadsorb_ = 8, SArevdes_ = 9, SAirrevdes_ = 10, SAflip_ = 11 };
enum { MSfront_ = 1, MSback_ = 2 };
constexpr double kVeryLarge = 1.0e+20;  // smoldyn.h:61

const char* shape_name(int ps) {
  static const char* n[6] = {"rect", "tri", "sph", "cyl", "hemi", "disk"};
  return (ps >= 0 && ps < 6) ? n[ps] : "none";
}

// the eight box-corner projections on v, in the reference's corner order
void corners(const double* v, const double* b1, const double* b2, double* o) {
  o[0] = v[0] * b1[0] + v[1] * b1[1] + v[2] * b1[2];
  o[1] = v[0] * b1[0] + v[1] * b1[1] + v[2] * b2[2];
  o[2] = v[0] * b1[0] + v[1] * b2[1] + v[2] * b1[2];
  o[3] = v[0] * b1[0] + v[1] * b2[1] + v[2] * b2[2];
  o[4] = v[0] * b2[0] + v[1] * b1[1] + v[2] * b1[2];
  o[5] = v[0] * b2[0] + v[1] * b1[1] + v[2] * b2[2];
  o[6] = v[0] * b2[0] + v[1] * b2[1] + v[2] * b1[2];
  o[7] = v[0] * b2[0] + v[1] * b2[1] + v[2] * b2[2];
}
bool all_below(const double* o, double c) {
  for (int k = 0; k < 8; k++)
    if (!(o[k] < c)) return false;
  return true;
}
bool all_above(const double* o, double c) {
  for (int k = 0; k < 8; k++)
    if (!(o[k] > c)) return false;
  return true;
}

// parallelogram r1, r2, r3 (r4 = r1 for a rectangle's call) against a box
bool rect_x_box(const double* r1, const double* r2, const double* r3, const double* r4, const double* b1,
                const double* b2) {
  double r5[3], r6[3], r7[3], o[8];
  for (int d = 0; d < 3; d++) {
    r5[d] = r2[d] + r4[d] - r1[d];
    r6[d] = r3[d] + r4[d] - r1[d];
    r7[d] = r2[d] + r3[d] - r1[d];
  }
  const double* rr[7] = {r1, r2, r3, r4, r5, r6, r7};
  for (int d = 0; d < 3; d++) {
    bool lo = true, hi = true;
    for (int k = 0; k < 7; k++) {
      if (!(rr[k][d] < b1[d])) lo = false;
      if (!(rr[k][d] > b2[d])) hi = false;
    }
    if (lo || hi) return false;
  }
  const double* ends[3] = {r2, r3, r4};
  for (int sdx = 0; sdx < 3; sdx++) {
    const double* ra = ends[sdx];
    double v[3] = {ra[0] - r1[0], ra[1] - r1[1], ra[2] - r1[2]};
    double cmp = v[0] * r1[0] + v[1] * r1[1] + v[2] * r1[2];
    corners(v, b1, b2, o);
    if (all_below(o, cmp)) return false;
    cmp = v[0] * ra[0] + v[1] * ra[1] + v[2] * ra[2];
    if (all_above(o, cmp)) return false;
  }
  return true;
}

// sphere surface against a box: only "box entirely inside" is excluded
// after the projection tests
bool sphere_x_box(const double* c, double rad, const double* b1, const double* b2) {
  for (int d = 0; d < 3; d++) {
    if (c[d] + rad < b1[d]) return false;
    if (c[d] - rad > b2[d]) return false;
  }
  double r2 = rad * rad, mn = 0, mx = 0;
  for (int k = 0; k < 8; k++) {
    const double* bx = (k & 4) ? b2 : b1;
    const double* by = (k & 2) ? b2 : b1;
    const double* bz = (k & 1) ? b2 : b1;
    double d2 = (bx[0] - c[0]) * (bx[0] - c[0]) + (by[1] - c[1]) * (by[1] - c[1]) + (bz[2] - c[2]) * (bz[2] - c[2]);
    if (k == 0) mn = mx = d2;
    else if (d2 < mn) mn = d2;
    else if (d2 > mx) mx = d2;
  }
  return !(r2 > mx);
}

// Geo_LineNormal3D (Geometry.c:103-160): unit vector from the line a-b to
// point, perpendicular to the line; returns the distance.  Degenerate lines
// give the unit vector to point (unit x when point == a); a point on the line
// gives (line_y, -line_x, 0) normalized, or unit x for a line along z.
double line_normal3(const double* a, const double* b, const double* point, double* ans) {
  const double eps = 100 * DBL_EPSILON;
  double line[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double dot = line[0] * line[0] + line[1] * line[1] + line[2] * line[2];
  if (dot < eps) {
    for (int d = 0; d < 3; d++) ans[d] = point[d] - a[d];
    dot = ans[0] * ans[0] + ans[1] * ans[1] + ans[2] * ans[2];
    if (dot < eps) {
      ans[0] = 1;
      ans[1] = ans[2] = 0;
      return 0;
    }
    dot = sqrt(dot);
    for (int d = 0; d < 3; d++) ans[d] /= dot;
    return dot;
  }
  dot = sqrt(dot);
  for (int d = 0; d < 3; d++) line[d] /= dot;
  for (int d = 0; d < 3; d++) ans[d] = point[d] - a[d];
  for (int pass = 0; pass < 2; pass++) {
    dot = ans[0] * line[0] + ans[1] * line[1] + ans[2] * line[2];
    for (int d = 0; d < 3; d++) ans[d] -= dot * line[d];
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
    for (int d = 0; d < 3; d++) ans[d] /= dot;
    return 0;
  }
  dot = sqrt(dot);
  for (int d = 0; d < 3; d++) ans[d] /= dot;
  return dot;
}

// Geo_TriXaabb3 (Geometry.c:1585-1605): bounding-box overlap, then the box
// corners not all strictly on one side of the triangle's plane
bool tri_x_box(const double* p1, const double* p2, const double* p3, const double* n, const double* b1,
               const double* b2) {
  for (int d = 0; d < 3; d++) {
    if (p1[d] < b1[d] && p2[d] < b1[d] && p3[d] < b1[d]) return false;
    if (p1[d] > b2[d] && p2[d] > b2[d] && p3[d] > b2[d]) return false;
  }
  double cmp = n[0] * p1[0] + n[1] * p1[1] + n[2] * p1[2], o[8];
  corners(n, b1, b2, o);
  return !all_below(o, cmp) && !all_above(o, cmp);
}

// Geo_LineXSphs (Geometry.c:1200-1221) in 2D without the nearest-point outputs
double line_x_circle(const double* p1, const double* p2, const double* c, double rad, double* cross2) {
  double a = 0, b = 0, cc = 0;
  for (int d = 0; d < 2; d++) {
    a += (p2[d] - p1[d]) * (p2[d] - p1[d]);
    b += (c[d] - p1[d]) * (p2[d] - p1[d]);
    cc += (p1[d] - c[d]) * (p1[d] - c[d]);
  }
  const double nrdist = sqrt(cc - b * b / a);
  b *= -2;
  cc -= rad * rad;
  double radical = b * b - 4 * a * cc;
  if (nrdist <= rad && radical < 0) radical = 0;
  radical = sqrt(radical);
  *cross2 = (-b + radical) / (2 * a);
  return (-b - radical) / (2 * a);
}

// Geo_LineXaabb (Geometry.c:1561-1582) for the infinite line through p1, p2
bool line_x_box(const double* p1, const double* p2, const double* b1, const double* b2) {
  double nearhi = -1.0, farlo = 2.0;
  for (int d = 0; d < 3; d++) {
    const double denom = p2[d] - p1[d];
    if (denom == 0) {
      if (p1[d] < b1[d] || p1[d] > b2[d]) return false;
    } else {
      const double a1 = (b1[d] - p1[d]) / denom, a2 = (b2[d] - p1[d]) / denom;
      if (a1 < a2) {
        if (a1 > nearhi) nearhi = a1;
        if (a2 < farlo) farlo = a2;
      } else {
        if (a2 > nearhi) nearhi = a2;
        if (a1 < farlo) farlo = a1;
      }
    }
  }
  return nearhi <= farlo;
}

// Geo_CylisXaabb3 (Geometry.c:1777-1916): infinite cylinder of axis p1-p2
// against a box, in the plane perpendicular to the axis (axes v1, v2):
// separating axes along the projected box edges, corner distances, the axis
// line through the box, and the projected box edges against the circle
bool cyl_inf_x_box(const double* p1, const double* p2, double rad, const double* b1, const double* b2) {
  static const int edge[12][2] = {{0, 1}, {1, 3}, {3, 2}, {2, 0}, {0, 4}, {1, 5},
                                  {2, 6}, {3, 7}, {4, 5}, {5, 7}, {7, 6}, {6, 4}};
  static const int axis[3] = {1, 2, 4};
  double v1[3], v2[3], v3[3], cent[2], ab[8][2], o1[8], o2[8], b[8];
  line_normal3(p1, p2, p1, v1);
  for (int d = 0; d < 3; d++) v3[d] = p2[d] - p1[d];
  double nrm = sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
  if (!nrm) return false;
  for (int d = 0; d < 3; d++) v3[d] /= nrm;
  v2[0] = v3[1] * v1[2] - v3[2] * v1[1];
  v2[1] = v3[2] * v1[0] - v3[0] * v1[2];
  v2[2] = v3[0] * v1[1] - v3[1] * v1[0];
  cent[0] = v1[0] * p1[0] + v1[1] * p1[1] + v1[2] * p1[2];
  cent[1] = v2[0] * p1[0] + v2[1] * p1[1] + v2[2] * p1[2];
  corners(v1, b1, b2, o1);
  corners(v2, b1, b2, o2);
  for (int k = 0; k < 8; k++) {
    ab[k][0] = o1[k];
    ab[k][1] = o2[k];
  }
  for (int a = 0; a < 3; a++) {
    double u[2] = {ab[axis[a]][0] - ab[0][0], ab[axis[a]][1] - ab[0][1]};
    nrm = sqrt(u[0] * u[0] + u[1] * u[1]);
    if (nrm > 0) {
      u[0] /= nrm;
      u[1] /= nrm;
      for (int k = 0; k < 8; k++) b[k] = u[0] * ab[k][0] + u[1] * ab[k][1];
      if (all_below(b, u[0] * cent[0] + u[1] * cent[1] - rad)) return false;
      if (all_above(b, u[0] * cent[0] + u[1] * cent[1] + rad)) return false;
    }
  }
  double mn = 0, mx = 0;
  for (int k = 0; k < 8; k++) {
    const double d2 = (ab[k][0] - cent[0]) * (ab[k][0] - cent[0]) + (ab[k][1] - cent[1]) * (ab[k][1] - cent[1]);
    if (k == 0) mn = mx = d2;
    else if (d2 < mn) mn = d2;
    else if (d2 > mx) mx = d2;
  }
  const double r2 = rad * rad;
  if (r2 > mx) return false;
  if (r2 >= mn) return true;
  if (line_x_box(p1, p2, b1, b2)) return true;
  for (int k = 0; k < 12; k++) {
    double c2;
    const double c1 = line_x_circle(ab[edge[k][0]], ab[edge[k][1]], cent, rad, &c2);
    if ((c1 >= 0 && c1 <= 1) || (c2 >= 0 && c2 <= 1)) return true;
  }
  return false;
}

// Geo_CylsXaabb3 (Geometry.c:1964-1970) with Geo_Cyl2Rect (:1137-1162)
bool cyl_x_box(const double* p1, const double* p2, double rad, const double* b1, const double* b2) {
  if (!cyl_inf_x_box(p1, p2, rad, b1, b2)) return false;
  double v1[3], v2[3], r[4][3];
  line_normal3(p1, p2, p1, v1);
  v2[0] = v1[1] * (p2[2] - p1[2]) - v1[2] * (p2[1] - p1[1]);
  v2[1] = v1[2] * (p2[0] - p1[0]) - v1[0] * (p2[2] - p1[2]);
  v2[2] = v1[0] * (p2[1] - p1[1]) - v1[1] * (p2[0] - p1[0]);
  const double nrm = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  for (int d = 0; d < 3; d++) v2[d] /= nrm;
  for (int d = 0; d < 3; d++) {
    r[0][d] = p1[d] + rad * (-v1[d] - v2[d]);
    r[1][d] = p1[d] + rad * (+v1[d] - v2[d]);
    r[2][d] = p1[d] + rad * (-v1[d] + v2[d]);
    r[3][d] = p2[d] + rad * (-v1[d] - v2[d]);
  }
  return rect_x_box(r[0], r[1], r[2], r[3], b1, b2);
}

// Geo_HemisXaabb3 (Geometry.c:1956-1962) with Geo_Hemis2Rect (:1113-1135)
bool hemi_x_box(const double* c, double rad, const double* out, const double* b1, const double* b2) {
  if (!sphere_x_box(c, rad, b1, b2)) return false;
  const double zero[3] = {0, 0, 0};
  double v1[3], v2[3], r[4][3];
  line_normal3(zero, out, zero, v1);
  v2[0] = v1[1] * out[2] - v1[2] * out[1];
  v2[1] = v1[2] * out[0] - v1[0] * out[2];
  v2[2] = v1[0] * out[1] - v1[1] * out[0];
  for (int d = 0; d < 3; d++) {
    r[0][d] = c[d] + rad * (-v1[d] - v2[d]);
    r[1][d] = c[d] + rad * (+v1[d] - v2[d]);
    r[2][d] = c[d] + rad * (-v1[d] + v2[d]);
    r[3][d] = c[d] + rad * (-v1[d] - v2[d] - out[d]);
  }
  return rect_x_box(r[0], r[1], r[2], r[3], b1, b2);
}

bool disk_x_box(const double* c, double rad, const double* n, const double* b1, const double* b2) {
  double dot = sqrt(n[1] * n[1] + n[2] * n[2]);
  if (c[0] - rad * dot > b2[0]) return false;
  if (c[0] + rad * dot < b1[0]) return false;
  dot = sqrt(n[0] * n[0] + n[2] * n[2]);
  if (c[1] - rad * dot > b2[1]) return false;
  if (c[1] + rad * dot < b1[1]) return false;
  dot = sqrt(n[0] * n[0] + n[1] * n[1]);
  if (c[2] - rad * dot > b2[2]) return false;
  if (c[2] + rad * dot < b1[2]) return false;
  double cmp = n[0] * c[0] + n[1] * c[1] + n[2] * c[2], o[8];
  corners(n, b1, b2, o);
  if (all_below(o, cmp)) return false;
  if (all_above(o, cmp)) return false;
  return true;
}

// srftristate2index (smolsurface.c:713-740)
void tristate2index(int ms, int ms1, int ms2, int* ms3p, int* facep, int* ms4p) {
  if (ms == MSnone_) ms = MSsoln_;
  int ms3 = (ms == MSsoln_ && (ms1 == MSsoln_ || ms1 == MSbsoln_)) ? MSsoln_ : (ms == MSsoln_ ? ms1 : ms);
  int face = ms1 == MSsoln_ ? PFfront_ : (ms1 == MSbsoln_ ? PFback_ : PFnone_);
  int ms4 = ms2;
  if (ms != MSsoln_ && ms1 != MSsoln_ && ms1 != MSbsoln_ && ms1 != ms) {
    ms3 = MSnone_;
    face = PFnone_;
    ms4 = MSnone_;
  }
  *ms3p = ms3;
  *facep = face;
  *ms4p = ms4;
}

// srfsamestate (smolsurface.c:640-658): whether ms2 is the state a molecule
// in ms1 at face1 stays in; *ms3p = that staying state
bool same_state(int ms1, int face1, int ms2, int* ms3p) {
  bool same;
  if (face1 == PFfront_ && ms2 == MSsoln_) same = true;
  else if (face1 == PFback_ && ms2 == MSbsoln_) same = true;
  else if (face1 == PFnone_ && ms1 == ms2) same = true;
  else same = false;
  if (ms3p) {
    if (face1 == PFfront_) *ms3p = MSsoln_;
    else if (face1 == PFback_) *ms3p = MSbsoln_;
    else *ms3p = (ms1 == MSsoln_) ? MSnone_ : ms1;
  }
  return same;
}

// srfreverseaction (smolsurface.c:661-710)
void reverse_action(int ms1, int face1, int ms2, int* ms3p, int* face2p, int* ms4p, int* actp) {
  int ms3, face2, ms4, act;
  if (ms1 == MSsoln_ && face1 == PFnone_) {
    ms3 = ms4 = MSnone_;
    face2 = PFnone_;
    act = SAnone_;
  } else if (ms1 == MSsoln_) {
    if (ms2 == MSsoln_ || ms2 == MSbsoln_) {
      ms3 = MSsoln_;
      face2 = (ms2 == MSsoln_) ? PFfront_ : PFback_;
      act = (face1 == face2) ? SAreflect_ : SAtrans_;
    } else {
      ms3 = ms2;
      face2 = PFnone_;
      act = SAadsorb_;
    }
    ms4 = (face1 == PFfront_) ? MSsoln_ : MSbsoln_;
  } else if (face1 == PFnone_) {
    if (ms2 == MSsoln_ || ms2 == MSbsoln_) {
      ms3 = MSsoln_;
      face2 = (ms2 == MSsoln_) ? PFfront_ : PFback_;
      act = SArevdes_;
    } else {
      ms3 = ms2;
      face2 = PFnone_;
      act = SAflip_;
    }
    ms4 = ms1;
  } else {
    if (ms2 == MSsoln_ || ms2 == MSbsoln_) {
      ms3 = ms1;
      face2 = (ms2 == MSsoln_) ? PFfront_ : PFback_;
      ms4 = (face1 == PFfront_) ? MSsoln_ : MSbsoln_;
      act = (face1 == face2) ? SAreflect_ : SAtrans_;
    } else {
      ms3 = ms2;
      face2 = PFboth_;
      ms4 = ms1;
      act = SAno_;
    }
  }
  if (ms3p) *ms3p = ms3;
  if (face2p) *face2p = face2;
  if (ms4p) *ms4p = ms4;
  if (actp) *actp = act;
}

// the rate branch of srfcalcprob (smolsurface.c:3097-3144): the probability
// that a molecule in state ms1 at face ends in ms2 when that transition is
// given by a rate, the reverse transition by raterev (0: none, -1: maximal).
// difc and difcrev are the solution diffusion coefficients of the species
// and of the one it becomes.  *st is set when the reference would leave the
// probability undefined (transmit_probs' codes).
double rate_prob(int ms1, int face, int ms2, double rate, double raterev, const SrfDetail& det, double difc,
                 double difcrev, double dt, int* st) {
  namespace sp = surfparam;
  bool to_soln = ms2 == MSsoln_ || ms2 == MSbsoln_;
  if (ms1 == MSsoln_ || face != PFnone_) {
    // at a face: a solution molecule, or a surface-bound one meeting a panel
    if ((face == PFfront_ && ms2 == MSsoln_) || (face == PFback_ && ms2 == MSbsoln_)) return 0;  // staying
    if (to_soln) {  // transmission
      if (raterev == 0) return sp::surfaceprob(rate, 0, dt, difc, nullptr, sp::kIrrTrans);
      double p = -2;
      *st = sp::transmit_probs(rate, raterev, difc, difcrev, dt, &p, nullptr);
      return p;
    }
    if (ms1 != MSsoln_) return sp::surfaceprob(rate, 0, dt, difc, nullptr, sp::kIrrAds);  // hop to another surface
    if (raterev == 0) return sp::surfaceprob(rate, 0, dt, difc, nullptr, sp::kIrrAds);  // adsorption
    return sp::surfaceprob(rate, raterev, dt, difc, nullptr, sp::kRevAds);
  }
  // surface-bound, leaving by desorption or a flip: first-order loss over the
  // sum of every exit rate when irreversible
  double sum = 0;
  for (int ms5 = 0; ms5 < kSrfMS1; ms5++)
    if (ms5 != ms1 && det.datasrc[ms5] == 1 && det.rate[ms5] >= 0) sum += det.rate[ms5];
  if (to_soln) {
    if (raterev == 0) return sp::surfaceprob(rate, sum, dt, difc, nullptr, sp::kIrrDes);
    return sp::surfaceprob(rate, raterev, dt, difc, nullptr, sp::kRevDes);
  }
  if (ms1 == ms2) return 0;
  if (raterev == 0) return sp::surfaceprob(rate, sum, dt, difc, nullptr, sp::kIrrFlip);
  return sp::surfaceprob(rate, raterev, dt, difc, nullptr, sp::kRevFlip);
}

std::unique_ptr<SrfDetail> new_detail(int species) {  // surfaceactionalloc
  std::unique_ptr<SrfDetail> a(new SrfDetail());
  memset(a.get(), 0, sizeof(SrfDetail));
  for (int ms = 0; ms < kSrfMS1; ms++) a->newspec[ms] = species;
  return a;
}

}  // namespace

int SurfaceSet::add_surface(const char* name) {
  std::unique_ptr<HostSurface> s(new HostSurface());
  s->name = name;
  for (int i = 0; i < kMaxSpecies; i++)
    for (int ms = 0; ms < kSrfMS; ms++) {  // surfacealloc defaults (smolsurface.c:1110-1118)
      s->action[i][ms][PFfront_] = s->action[i][ms][PFback_] = SAtrans_;
      s->action[i][ms][PFnone_] = SAno_;
    }
  srf.push_back(std::move(s));
  return (int)srf.size() - 1;
}

int SurfaceSet::find_surface(const char* name) const {
  for (size_t s = 0; s < srf.size(); s++)
    if (srf[s]->name == name) return (int)s;
  return -1;
}

// surfsetaction (smolsurface.c:2158-2276) with no species change
int SurfaceSet::set_action(int s, int ident, int nspecies, int ms, int face, int act) {
  HostSurface& S = *srf[s];
  int mslo = ms, mshi = ms;
  if (ms == MSall_) {
    mslo = 0;
    mshi = kSrfMS - 1;
  }
  int ilo = ident, ihi = ident;
  if (ident == -5) {
    ilo = 1;
    ihi = nspecies - 1;
  }
  for (int i = ilo; i <= ihi; i++)
    for (int m = mslo; m <= mshi; m++) {
      if (face == PFfront_ || face == PFboth_) S.action[i][m][PFfront_] = act;
      if (face == PFback_ || face == PFboth_) S.action[i][m][PFback_] = act;
      if (face == PFnone_) S.action[i][m][PFnone_] = act;
    }
  return 0;
}

// surfsetrate (smolsurface.c:2278-2329)
int SurfaceSet::set_rate(int s, int ident, int nspecies, int ms, int ms1, int ms2, int newident, double value,
                         int which) {
  HostSurface& S = *srf[s];
  int ms3, face, ms4;
  tristate2index(ms, ms1, ms2, &ms3, &face, &ms4);
  if (ms3 == MSnone_) return 3;
  int ilo = ident, ihi = ident;
  if (ident == -5) {
    ilo = 1;
    ihi = nspecies - 1;
  }
  for (int i = ilo; i <= ihi; i++) {
    if (!S.det[i][ms3][face]) S.det[i][ms3][face] = new_detail(i);
    SrfDetail& a = *S.det[i][ms3][face];
    S.action[i][ms3][face] = SAmult_;
    if (which == 1) {
      a.rate[ms4] = value;
      a.datasrc[ms4] = 1;
    } else {
      a.prob[ms4] = value;
      a.datasrc[ms4] = 2;
    }
    a.newspec[ms4] = (newident == -5) ? i : newident;
  }
  return 0;
}

// surfaddpanel (smolsurface.c:2432-2640) for 3D panels of every shape; a
// panel of an existing name replaces it
int SurfaceSet::add_panel(int s, int ps, const char* name, const char* axis, const double* params) {
  HostPanel p;
  p.ps = ps;
  p.srf = s;
  double(*point)[3] = p.point;
  double* front = p.front;
  if (ps == 0) {
    if (!axis || !(axis[0] == '+' || axis[0] == '-')) return -3;
    int pdim = axis[1] == '0' || axis[1] == 'x' ? 0 : axis[1] == '1' || axis[1] == 'y' ? 1 : axis[1] == '2' || axis[1] == 'z' ? 2 : 4;
    if (pdim > 2) return -3;
    int pt = 1;
    if (axis[0] == '-') pt *= -1;
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
    front[0] = axis[0] == '+' ? 1 : -1;
    front[1] = pdim;
    int axis01 = pdim == 0 ? (pt == 1 ? 1 : 2) : pdim == 1 ? (pt == 1 ? 2 : 0) : (pt == 1 ? 0 : 1);
    front[2] = axis01;
  } else if (ps == 2) {
    point[0][0] = params[0];
    point[0][1] = params[1];
    point[0][2] = params[2];
    point[1][0] = fabs(params[3]);
    point[1][1] = params[4];
    point[1][2] = params[5];
    if (point[1][1] <= 0 || point[1][2] <= 0) return -4;
    front[0] = params[3] > 0 ? 1 : (params[3] < 0 ? -1 : 0);
  } else if (ps == 1) {
    // vertices, unit normal (Geo_TriNormal), outward unit edge normals
    // (Geo_UnitCross of each edge with the normal), smolsurface.c:2521-2534
    for (int k = 0; k < 3; k++)
      for (int d = 0; d < 3; d++) point[k][d] = params[3 * k + d];
    double e1[3], e2[3];
    for (int d = 0; d < 3; d++) {
      e1[d] = point[1][d] - point[0][d];
      e2[d] = point[2][d] - point[1][d];
    }
    front[0] = e1[1] * e2[2] - e1[2] * e2[1];
    front[1] = -e1[0] * e2[2] + e1[2] * e2[0];
    front[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double area = sqrt(front[0] * front[0] + front[1] * front[1] + front[2] * front[2]);
    if (area > 100 * DBL_EPSILON) {
      for (int d = 0; d < 3; d++) front[d] /= area;
    } else {  // Geo_LineNormal of the first edge, in the xy plane
      const double len = sqrt(e1[0] * e1[0] + e1[1] * e1[1]);
      if (len > 0) {
        const double leninv = 1.0 / len;
        front[0] = leninv * e1[1];
        front[1] = -leninv * e1[0];
      } else {
        front[0] = 1;
        front[1] = 0;
      }
      front[2] = 0;
    }
    for (int k = 0; k < 3; k++) {
      const double* a = point[k];
      const double* b = point[(k + 1) % 3];
      const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double c[3] = {u[1] * front[2] - u[2] * front[1], u[2] * front[0] - u[0] * front[2],
                           u[0] * front[1] - u[1] * front[0]};
      const double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      for (int d = 0; d < 3; d++) point[3 + k][d] = len > 100 * DBL_EPSILON ? c[d] / len : 0;
    }
  } else if (ps == 3) {
    for (int k = 0; k < 2; k++)
      for (int d = 0; d < 3; d++) point[k][d] = params[3 * k + d];
    if (point[0][0] == point[1][0] && point[0][1] == point[1][1] && point[0][2] == point[1][2]) return -5;
    point[2][0] = fabs(params[6]);
    point[2][1] = params[7];
    point[2][2] = params[8];
    if (point[2][1] <= 0 || point[2][2] <= 0) return -4;
    front[2] = params[6] > 0 ? 1 : (params[6] < 0 ? -1 : 0);
    double len = 0;
    for (int d = 0; d < 3; d++) len += (point[1][d] - point[0][d]) * (point[1][d] - point[0][d]);
    const double lengthinv = 1.0 / sqrt(len);
    for (int d = 0; d < 3; d++) {
      point[3][d] = -(point[1][d] - point[0][d]) * lengthinv;
      point[4][d] = (point[1][d] - point[0][d]) * lengthinv;
    }
  } else if (ps == 4) {
    for (int d = 0; d < 3; d++) point[0][d] = params[d];
    point[1][0] = fabs(params[3]);
    point[1][1] = params[7];
    point[1][2] = params[8];
    if (point[1][1] <= 0 || point[1][2] <= 0) return -4;
    double len = 0;  // normalizeVD (Rn.c:500-508)
    for (int d = 0; d < 3; d++) len += params[4 + d] * params[4 + d];
    if (!len) return -6;
    len = sqrt(len);
    for (int d = 0; d < 3; d++) point[2][d] = params[4 + d] / len;
    front[0] = params[3] > 0 ? 1 : (params[3] < 0 ? -1 : 0);
  } else if (ps == 5) {
    point[0][0] = params[0];
    point[0][1] = params[1];
    point[0][2] = params[2];
    point[1][0] = params[3];
    if (point[1][0] <= 0) return -7;
    point[1][1] = params[7];
    if (point[1][1] <= 0) return -4;
    double len = sqrt(params[4] * params[4] + params[5] * params[5] + params[6] * params[6]);  // normalizeVD
    if (len <= 0) return -8;
    front[0] = params[4] / len;
    front[1] = params[5] / len;
    front[2] = params[6] / len;
  } else {
    return -2;
  }
  HostSurface& S = *srf[s];
  int found = -1;
  if (name && name[0]) {
    for (int q = 0; q < 6 && found < 0; q++)
      for (int id : S.panels[q])
        if (pnl[id].name == name) {
          if (q != ps) return -9;
          found = id;
        }
  }
  if (found >= 0) {
    p.name = pnl[found].name;
    p.jumpp[0] = pnl[found].jumpp[0];
    p.jumpp[1] = pnl[found].jumpp[1];
    p.jumpf[0] = pnl[found].jumpf[0];
    p.jumpf[1] = pnl[found].jumpf[1];
    pnl[found] = p;
    return found;
  }
  char def[64];
  snprintf(def, sizeof(def), "%s%d", shape_name(ps), S.count[ps]);
  p.name = (name && name[0]) ? name : def;
  S.count[ps]++;
  pnl.push_back(p);
  S.panels[ps].push_back((int)pnl.size() - 1);
  return (int)pnl.size() - 1;
}

int SurfaceSet::find_panel(int s, const char* name) const {
  for (int q = 0; q < 6; q++)
    for (int id : srf[s]->panels[q])
      if (pnl[id].name == name) return id;
  return -1;
}

int SurfaceSet::set_jump(int p1, int f1, int p2, int f2, int bidirectional) {
  if (!(f1 == PFfront_ || f1 == PFback_) || !(f2 == PFfront_ || f2 == PFback_)) return 2;
  pnl[p1].jumpp[f1] = p2;
  pnl[p1].jumpf[f1] = f2;
  if (bidirectional) {
    pnl[p2].jumpp[f2] = p1;
    pnl[p2].jumpf[f2] = f1;
  }
  return 0;
}

int SurfaceSet::build(const Geometry& g, int nspecies, const double (*difc_ms)[kSrfMS], double dt,
                      const int (*lookup_ms)[kSrfMS], const bool* mobile_list, int nlist, SurfaceTables& out,
                      char* msg, int msglen) const {
  int nsrf = (int)srf.size();
  if (nsrf > kMaxSurf) {
    snprintf(msg, msglen, "more than %d surfaces", kMaxSurf);
    return 1;
  }
  // srfmollist SMLdiffuse (surfupdatelists, smolsurface.c:3830-3845): lists
  // of mobile molecules when any surface has an action other than
  // transmission
  bool check = false;
  for (int s = 0; s < nsrf && !check; s++)
    for (int i = 1; i < nspecies && !check; i++)
      for (int ms = 0; ms < kSrfMS && !check; ms++)
        for (int f = 0; f <= PFnone_ && !check; f++) {
          int act = srf[s]->action[i][ms][f];
          if (!(act == SAno_ || act == SAtrans_) || srf[s]->det[i][ms][f]) check = true;
        }
  for (int ll = 0; ll < kMaxList; ll++) out.listdiffuse[ll] = (ll < nlist && check && mobile_list[ll]) ? 1 : 0;

  // panels
  out.pnl.clear();
  for (const HostPanel& p : pnl) {
    DevPanel d;
    memset(&d, 0, sizeof(d));
    d.ps = p.ps;
    d.srf = p.srf;
    for (int k = 0; k < 2; k++) {
      d.jumpp[k] = p.jumpp[k];
      d.jumpf[k] = p.jumpf[k];
    }
    for (int a = 0; a < kPanelPoints; a++)
      for (int b = 0; b < 3; b++) d.point[a][b] = p.point[a][b];
    for (int b = 0; b < 3; b++) d.front[b] = p.front[b];
    out.pnl.push_back(d);
  }

  // probabilities of the rate-based actions (srfcalcprob, smolsurface.c:
  // 3070-3146, for the cases restated here: given probabilities, and
  // desorption or state flips by rate without a reverse rate), their
  // normalization (smolsurface.c:3758-3772) and cumulative sums, and the
  // reverse probabilities that tell reversible from irreversible desorption
  struct Probs {
    double prob[kSrfMS1], cum[kSrfMS1], rev[kSrfMS1];
  };
  std::vector<Probs> pr((size_t)nsrf * kMaxSpecies * kSrfMS * 3);
  auto P = [&](int s, int i, int ms, int f) -> Probs& { return pr[(((size_t)s * kMaxSpecies + i) * kSrfMS + ms) * 3 + f]; };
  for (int s = 0; s < nsrf; s++) {
    const HostSurface& S = *srf[s];
    for (int i = 1; i < nspecies; i++)
      for (int ms = 0; ms < kSrfMS; ms++)
        for (int f = 0; f < 3; f++) {
          const SrfDetail* det = S.det[i][ms][f].get();
          if (S.action[i][ms][f] != SAmult_ || !det) continue;
          Probs& q = P(s, i, ms, f);
          for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
            double p = 0;
            if (ms == MSsoln_ && f == PFnone_) p = 0;
            else if (det->datasrc[ms2] == 2) p = det->prob[ms2];
            else if (det->datasrc[ms2] == 1 && det->rate[ms2] != 0) {
              double rate = det->rate[ms2];
              if (rate < 0 && rate != -1) {
                snprintf(msg, msglen, "surface %s: negative rate", S.name.c_str());
                return 3;
              }
              // reverse rate (srfreverseaction, smolsurface.c:661-710)
              int ms3, face2, ms4, actrev;
              reverse_action(ms, f, ms2, &ms3, &face2, &ms4, &actrev);
              int i2 = det->newspec[ms2];
              double raterev = 0;
              if (i2 && actrev != SAno_ && ms3 < kSrfMS && face2 < 3) {
                if (srf[s]->action[i2][ms3][face2] == SAmult_) {
                  const SrfDetail* b2 = srf[s]->det[i2][ms3][face2].get();
                  if (b2 && b2->datasrc[ms4] == 1) raterev = b2->rate[ms4];
                } else if (srf[s]->action[i2][ms3][face2] == actrev) {
                  raterev = -1;
                }
              }
              if (raterev < 0 && raterev != -1) raterev = 0;
              int st = 0;
              p = rate_prob(ms, f, ms2, rate, raterev, *det, difc_ms[i][MSsoln_], difc_ms[i2][MSsoln_], dt, &st);
              if (st) {
                snprintf(msg, msglen,
                         st == 3 ? "surface %s: a two-sided transmission with one side at the maximal rate is "
                                   "converted by the reference with a randomized search, which is not restated"
                                 : "surface %s: transmission probability undefined (zero diffusion coefficient)",
                         S.name.c_str());
                return 3;
              }
            }
            q.prob[ms2] = p;
          }
          // normalization: the probabilities of leaving the state are scaled
          // down when they exceed 1, else the staying probability is the rest
          double sum = 0;
          for (int ms2 = 0; ms2 < kSrfMS1; ms2++)
            if (!same_state(ms, f, ms2, nullptr)) sum += q.prob[ms2];
          if (sum > 1.0) {
            q.prob[ms] = 0;
            for (int ms2 = 0; ms2 < kSrfMS1; ms2++) q.prob[ms2] /= sum;
          } else {
            int ms3;
            same_state(ms, f, MSsoln_, &ms3);
            if (ms3 < kSrfMS1) q.prob[ms3] = 1.0 - sum;
          }
          sum = 0;
          for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
            sum += q.prob[ms2];
            q.cum[ms2] = sum;
          }
          // the reference's parameter check rejects these (checksurfaceparams,
          // smolsurface.c:1940-1943); a rate-given reverse desorption at the
          // maximal rate gives one (SurfaceParam.c:170-172)
          for (int ms2 = 0; ms2 < kSrfMS1; ms2++)
            if (!(q.prob[ms2] >= 0 && q.prob[ms2] <= 1)) {
              snprintf(msg, msglen, "surface %s: interaction probability %g of species %d is outside [0, 1]",
                       S.name.c_str(), q.prob[ms2], i);
              return 3;
            }
        }
    for (int i = 1; i < nspecies; i++)
      for (int ms = 0; ms < kSrfMS; ms++)
        for (int f = 0; f < 3; f++) {
          const SrfDetail* det = S.det[i][ms][f].get();
          if (S.action[i][ms][f] != SAmult_ || !det) continue;
          for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
            int ms3, face2, ms4;
            reverse_action(ms, f, ms2, &ms3, &face2, &ms4, nullptr);
            int i2 = det->newspec[ms2];
            if (face2 != PFboth_ && ms3 < kSrfMS && face2 < 3 && S.det[i][ms3][face2] && S.det[i2][ms3][face2])
              P(s, i, ms, f).rev[ms2] = P(s, i2, ms3, face2).prob[ms4];
          }
        }
  }

  // device action tables: solution-phase molecules at the front and back
  // faces; surface-bound molecules (front or back state) with no face
  out.act.assign((size_t)nsrf * kMaxSpecies * 2, DevSrfAct());
  out.actb.assign((size_t)nsrf * kMaxSpecies * 2, DevSrfAct());
  for (int s = 0; s < nsrf; s++) {
    const HostSurface& S = *srf[s];
    for (int i = 1; i < nspecies; i++)
      for (int f = 0; f < 2; f++) {
        DevSrfAct& a = out.act[((size_t)s * kMaxSpecies + i) * 2 + f];
        memset(&a, 0, sizeof(a));
        int act = S.action[i][MSsoln_][f];
        a.act = act;
        a.newspec_fixed = i;
        const SrfDetail* det = S.det[i][MSsoln_][f].get();
        if (act != SAmult_) {
          if (det && det->datasrc[MSsoln_] == 3) a.newspec_fixed = det->newspec[MSsoln_];
          if (act == SAport_) {
            snprintf(msg, msglen, "surface %s: port actions are not implemented on the device path",
                     S.name.c_str());
            return 2;
          }
          continue;
        }
        if (!det) {
          snprintf(msg, msglen, "surface %s: rate-based action without rates", S.name.c_str());
          return 2;
        }
        const Probs& q = P(s, i, MSsoln_, f);
        for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
          a.cumprob[ms2] = q.cum[ms2];
          a.newspec[ms2] = det->newspec[ms2];
          int o;
          if (det->newspec[ms2] == 0) o = SAabsorb_;
          else if (f == PFfront_) o = ms2 == MSsoln_ ? SAreflect_ : ms2 == MSbsoln_ ? SAtrans_ : SAadsorb_;
          else o = ms2 == MSsoln_ ? SAtrans_ : ms2 == MSbsoln_ ? SAreflect_ : SAadsorb_;
          a.outcome[ms2] = o;
          bool possible = q.prob[ms2] > 0;
          if (possible && o == SAadsorb_) {
            int i2 = det->newspec[ms2];
            if (ms2 != MSfront_ && ms2 != MSback_) {
              snprintf(msg, msglen, "surface %s: adsorption to the up/down states is not implemented",
                       S.name.c_str());
              return 2;
            }
            if (difc_ms[i2][ms2] > 0) {
              snprintf(msg, msglen, "surface %s: diffusion of surface-bound molecules is not implemented",
                       S.name.c_str());
              return 2;
            }
          }
        }
      }
    for (int i = 1; i < nspecies; i++)
      for (int b = 0; b < 2; b++) {
        const int ms = b == 0 ? MSfront_ : MSback_;
        DevSrfAct& a = out.actb[((size_t)s * kMaxSpecies + i) * 2 + b];
        memset(&a, 0, sizeof(a));
        int act = S.action[i][ms][PFnone_];
        a.act = act == SAmult_ ? SAmult_ : SAno_;
        a.newspec_fixed = i;
        if (act != SAmult_) {
          if (act != SAno_) {
            snprintf(msg, msglen, "surface %s: fixed actions of surface-bound molecules are not implemented",
                     S.name.c_str());
            return 2;
          }
          continue;
        }
        const SrfDetail* det = S.det[i][ms][PFnone_].get();
        if (!det) continue;
        const Probs& q = P(s, i, ms, PFnone_);
        for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
          a.cumprob[ms2] = q.cum[ms2];
          a.newspec[ms2] = det->newspec[ms2];
          int o;
          if (det->newspec[ms2] == 0) o = SAabsorb_;
          else if (ms2 == ms) o = SAno_;
          else if (ms2 == MSsoln_ || ms2 == MSbsoln_) o = q.rev[ms2] > 0 ? SArevdes_ : SAirrevdes_;
          else o = SAflip_;
          a.outcome[ms2] = o;
          // flips between the front and back states run on the device;
          // flips to the up and down states do not
          if (q.prob[ms2] > 0 && o == SAflip_ && ms2 != MSfront_ && ms2 != MSback_) {
            snprintf(msg, msglen, "surface %s: flips to the up or down state are not implemented on the device path",
                     S.name.c_str());
            return 2;
          }
        }
      }
  }
  // lists holding surface-bound states (srfmollist SMLsrfbound) and the
  // list of each species in the front and back states
  for (int ll = 0; ll < kMaxList; ll++) out.listbound[ll] = 0;
  for (int i = 0; i < kMaxSpecies; i++) out.blist[i][0] = out.blist[i][1] = -1;
  for (int i = 1; i < nspecies; i++)
    for (int ms = 1; ms < kSrfMS; ms++) {
      int ll = lookup_ms[i][ms];
      if (ll >= 0 && ll < kMaxList) out.listbound[ll] = 1;
      if (ms == MSfront_ || ms == MSback_) out.blist[i][ms - 1] = ll;
    }

  // box panel lists (boxesupdateparams, smolboxes.c:652-671 with panelinbox
  // :95-152); edge boxes extend to +-1e20 as the reference's do
  out.bpstart.assign((size_t)g.nbox + 1, 0);
  out.bplist.clear();
  int n = 0;
  for (int b = 0; b < g.nbox; b++) {
    int idx[3] = {b / (g.side[1] * g.side[2]), (b / g.side[2]) % g.side[1], b % g.side[2]};
    double v1[3], v2[3];
    for (int d = 0; d < 3; d++) {
      v1[d] = g.bmin[d] + idx[d] * g.bsize[d];
      v2[d] = g.bmin[d] + (idx[d] + 1) * g.bsize[d];
      if (idx[d] == 0) v1[d] = -kVeryLarge;
      if (idx[d] == g.side[d] - 1) v2[d] = kVeryLarge;
    }
    out.bpstart[b] = n;
    for (int s = 0; s < nsrf; s++)
      for (int ps = 0; ps < 6; ps++)
        for (int id : srf[s]->panels[ps]) {
          const HostPanel& p = pnl[id];
          bool in;
          if (ps == 0) in = rect_x_box(p.point[0], p.point[1], p.point[3], p.point[0], v1, v2);
          else if (ps == 1) in = tri_x_box(p.point[0], p.point[1], p.point[2], p.front, v1, v2);
          else if (ps == 2) in = sphere_x_box(p.point[0], p.point[1][0], v1, v2);
          else if (ps == 3) in = cyl_x_box(p.point[0], p.point[1], p.point[2][0], v1, v2);
          else if (ps == 4) in = hemi_x_box(p.point[0], p.point[1][0], p.point[2], v1, v2);
          else in = disk_x_box(p.point[0], p.point[1][0], p.front, v1, v2);
          if (in) {
            out.bplist.push_back(id);
            n++;
          }
        }
  }
  out.bpstart[g.nbox] = n;
  return 0;
}

}  // namespace smolamd
