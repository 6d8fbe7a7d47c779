// Device restatement of the reference's molecule-surface path for
// solution-phase molecules in 3D (source/Smoldyn/smolsurface.c,
// smolboxes.c:846-892 line2nextbox, source/libSteve/Geometry.c).  Panels:
// rectangle, triangle, sphere, cylinder, hemisphere, disk.  Actions: reflect,
// transmit, jump, and rate-based
// actions whose outcomes are those; absorption and adsorption are reported to
// the caller (the kernels raise kErrSurface for them).
//
// The arithmetic follows the reference operation by operation (the library
// is compiled with -ffp-contract=off), so collisions give the same FP64
// positions as the CPU path.  Included by engine.hip only.
#ifndef SMOLDYN_AMD_SURFACE_DEV_H
#define SMOLDYN_AMD_SURFACE_DEV_H

namespace smolamd {

constexpr double kVeryClose = 1.0e-12;                // smoldyn.h:60
constexpr double kDblEpsilon = 2.220446049250313e-16;  // DBL_EPSILON
enum { kFaceFront = 0, kFaceBack = 1, kFaceNone = 2 };

// Geo_LineNormal3D (Geometry.c:103-160): unit vector from the line a-b to
// point, perpendicular to the line; returns the distance.  Degenerate lines
// give the unit vector to point (unit x when point == a); a point on the line
// gives (line_y, -line_x, 0) normalized, or unit x for a line along z.
__device__ __forceinline__ double srf_linenormal3(const double* a, const double* b, const double* point, double* ans) {
  const double eps = 100 * kDblEpsilon;
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

// Geo_PtInSlab (Geometry.c:495-505): test between the planes through a and
// b perpendicular to a-b
__device__ __forceinline__ bool srf_ptinslab(const double* a, const double* b, const double* t) {
  double dot = 0;
  for (int d = 0; d < 3; d++) dot += (t[d] - a[d]) * (b[d] - a[d]);
  if (dot < 0) return false;
  dot = 0;
  for (int d = 0; d < 3; d++) dot += (t[d] - b[d]) * (a[d] - b[d]);
  return !(dot < 0);
}

// Geo_PtInTriangle2 (Geometry.c:474-492): on no edge's outer side
__device__ __forceinline__ bool srf_ptintriangle(const DevPanel& pn, const double* t) {
  for (int k = 0; k < 3; k++) {
    const double* p = pn.point[k];
    const double* e = pn.point[k + 3];
    if ((t[0] - p[0]) * e[0] + (t[1] - p[1]) * e[1] + (t[2] - p[2]) * e[2] > 0) return false;
  }
  return true;
}

// panelside (smolsurface.c:3930-3963)
template <bool kCv>
__device__ __forceinline__ int srf_panelside(const DevPanel& pn, const double* pt, double* distp, int strict) {
  double dist = 0;
  if (pn.ps == kPsRect) {
    int a = (int)pn.front[1];
    dist = pn.front[0] * (pt[a] - pn.point[0][a]);
  } else if (pn.ps == kPsDisk || (kCv && pn.ps == kPsTri)) {
#pragma unroll
    for (int d = 0; d < 3; d++) dist += (pt[d] - pn.point[0][d]) * pn.front[d];
  } else if (kCv && pn.ps == kPsCyl) {
    double cylnorm[3];
    dist = srf_linenormal3(pn.point[0], pn.point[1], pt, cylnorm);
    dist = pn.front[2] * (dist - pn.point[2][0]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; d++) dist += (pt[d] - pn.point[0][d]) * (pt[d] - pn.point[0][d]);
    dist = pn.front[0] * (sqrt(dist) - pn.point[1][0]);
  }
  if (distp) *distp = dist;
  if (dist > 0) return kFaceFront;
  if (!strict || dist < 0) return kFaceBack;
  return kFaceNone;
}

// Geo_SphereNormal (Geometry.c:199-213)
__device__ __forceinline__ void srf_spherenormal(const double* cent, const double* pt, int front, double* ans) {
  double dist = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    ans[d] = front * (pt[d] - cent[d]);
    dist += ans[d] * ans[d];
  }
  if (dist > 0) {
    dist = sqrt(dist);
#pragma unroll
    for (int d = 0; d < 3; d++) ans[d] /= dist;
  } else {
    ans[0] = 1;
    ans[1] = 0;
    ans[2] = 0;
  }
}

// Geo_LineXSphs (Geometry.c:1200-1221)
__device__ __forceinline__ double srf_linexsphs(const double* pt1, const double* pt2, const double* cent, double rad,
                                                double* cross2p, double* nrdistp, double* nrposp) {
  double a = 0, b = 0, c = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a += (pt2[d] - pt1[d]) * (pt2[d] - pt1[d]);
    b += (cent[d] - pt1[d]) * (pt2[d] - pt1[d]);
    c += (pt1[d] - cent[d]) * (pt1[d] - cent[d]);
  }
  double nrdist = sqrt(c - b * b / a);
  *nrdistp = nrdist;
  *nrposp = b / a;
  b *= -2;
  c -= rad * rad;
  double radical = b * b - 4 * a * c;
  if (nrdist <= rad && radical < 0) radical = 0;
  radical = sqrt(radical);
  *cross2p = (-b + radical) / (2 * a);
  return (-b - radical) / (2 * a);
}

// lineXpanel (smolsurface.c:3999-4101), 3D panels
template <bool kCv>
__device__ __forceinline__ int srf_linexpanel(const DevPanel& pn, const double* pt1, const double* pt2,
                                              double* crsspt, int* face1p, int* face2p, double* crossp,
                                              double* cross2p) {
  double dist1, dist2, cross = -1, cross2 = -1;
  int face1 = srf_panelside<kCv>(pn, pt1, &dist1, 0);
  int face2 = srf_panelside<kCv>(pn, pt2, &dist2, 0);
  int intsct;
  if (pn.ps == kPsRect) {
    if (face1 == face2) return 0;
    cross = dist1 / (dist1 - dist2);
#pragma unroll
    for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    int d = (int)pn.front[2];
    intsct = ((pn.point[0][d] <= crsspt[d] && crsspt[d] <= pn.point[1][d]) ||
              (pn.point[1][d] <= crsspt[d] && crsspt[d] <= pn.point[0][d]));
    d = (d + 1) % 3;
    if (d == (int)pn.front[1]) d = (d + 1) % 3;
    intsct = intsct && ((pn.point[1][d] <= crsspt[d] && crsspt[d] <= pn.point[2][d]) ||
                        (pn.point[2][d] <= crsspt[d] && crsspt[d] <= pn.point[1][d]));
  } else if (pn.ps == kPsSph || (kCv && pn.ps == kPsHemi)) {
    int facein = pn.front[0] > 0 ? kFaceBack : kFaceFront;
    if (face1 == facein && face2 == facein) return 0;
    double nrdist, nrpos;
    cross = srf_linexsphs(pt1, pt2, pn.point[0], pn.point[1][0], &cross2, &nrdist, &nrpos);
    if (face1 == face2 && (!(nrdist <= pn.point[1][0]) || !(nrpos >= 0) || !(nrpos <= 1))) return 0;
    if (face1 == facein) cross = cross2;
#pragma unroll
    for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    intsct = 1;
    if (kCv && pn.ps == kPsHemi) {  // the crossing must be on the dome side of the opening
      double dot = 0;
      for (int d = 0; d < 3; d++) dot += (crsspt[d] - pn.point[0][d]) * pn.point[2][d];
      intsct = (dot <= 0);
      if (!intsct && face1 == face2) {
        cross = cross2;
        face1 = (face1 == kFaceFront) ? kFaceBack : kFaceFront;
        for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
        dot = 0;
        for (int d = 0; d < 3; d++) dot += (crsspt[d] - pn.point[0][d]) * pn.point[2][d];
        intsct = (dot <= 0);
      }
    }
  } else if (kCv && pn.ps == kPsCyl) {
    // Geo_LineXCyls (Geometry.c:1248-1258): the segment's perpendicular
    // offsets from the axis against a circle about the origin
    int facein = (int)pn.front[2] > 0 ? kFaceBack : kFaceFront;
    if (face1 == facein && face1 == face2) return 0;
    double n1[3], n2[3], zero[3] = {0, 0, 0}, nrdist, nrpos;
    const double r1 = srf_linenormal3(pn.point[0], pn.point[1], pt1, n1);
    const double r2 = srf_linenormal3(pn.point[0], pn.point[1], pt2, n2);
    for (int d = 0; d < 3; d++) {
      n1[d] *= r1;
      n2[d] *= r2;
    }
    cross = srf_linexsphs(n1, n2, zero, pn.point[2][0], &cross2, &nrdist, &nrpos);
    if (face1 == face2 && (!(nrdist <= pn.point[2][0]) || !(nrpos >= 0) || !(nrpos <= 1))) return 0;
    if (face1 == facein) cross = cross2;
    for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    intsct = srf_ptinslab(pn.point[0], pn.point[1], crsspt);
    if (!intsct && face1 == face2) {
      cross = cross2;
      face1 = (face1 == kFaceFront) ? kFaceBack : kFaceFront;
      for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
      intsct = srf_ptinslab(pn.point[0], pn.point[1], crsspt);
    }
  } else if (kCv && pn.ps == kPsTri) {
    if (face1 == face2) return 0;
    cross = dist1 / (dist1 - dist2);
    for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    intsct = srf_ptintriangle(pn, crsspt);
  } else {
    if (face1 == face2) return 0;
    cross = dist1 / (dist1 - dist2);
#pragma unroll
    for (int d = 0; d < 3; d++) crsspt[d] = pt1[d] + cross * (pt2[d] - pt1[d]);
    double dot = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) dot += (crsspt[d] - pn.point[0][d]) * (crsspt[d] - pn.point[0][d]);
    intsct = (dot <= pn.point[1][0] * pn.point[1][0]);
  }
  *face1p = face1;
  if (face2p) *face2p = face2;
  *crossp = cross;
  if (cross2p) *cross2p = cross2;
  return intsct;
}

// getpanelnormal (smolsurface.c:4430-4456)
template <bool kCv>
__device__ __forceinline__ void srf_getpanelnormal(const DevPanel& pn, const double* pt, double* norm) {
  if (pn.ps == kPsRect) {
    norm[0] = norm[1] = norm[2] = 0;
    norm[(int)pn.front[1]] = pn.front[0];
  } else if (pn.ps == kPsDisk || (kCv && pn.ps == kPsTri)) {
    norm[0] = pn.front[0];
    norm[1] = pn.front[1];
    norm[2] = pn.front[2];
  } else if (kCv && pn.ps == kPsCyl) {
    srf_linenormal3(pn.point[0], pn.point[1], pt, norm);
    if (pn.front[2] == -1)
      for (int d = 0; d < 3; d++) norm[d] *= -1;
  } else {
    srf_spherenormal(pn.point[0], pt, (int)pn.front[0], norm);
  }
}

// fixpt2panel (smolsurface.c:4460-4474): onto the panel, then nudged by
// doubling multiples of DBL_EPSILON until on the requested side
template <bool kCv>
__device__ __forceinline__ void srf_fixpt2panel(const DevPanel& pn, double* pt, int face, double epsilon) {
  double dist, norm[3];
  int faceat = srf_panelside<kCv>(pn, pt, &dist, 1);
  if ((faceat == face || face == kFaceNone) && fabs(dist) <= epsilon) return;
  srf_getpanelnormal<kCv>(pn, pt, norm);
#pragma unroll
  for (int d = 0; d < 3; d++) pt[d] -= dist * norm[d];
  if (face == kFaceNone) return;
  int sign = (face == kFaceFront) ? 1 : -1;
  // each pass doubles the step; 1100 doublings exceed any finite coordinate
  for (double factor = 1.0; face != srf_panelside<kCv>(pn, pt, nullptr, 1) && factor < 1e300; factor *= 2.0) {
#pragma unroll
    for (int d = 0; d < 3; d++) pt[d] += sign * factor * kDblEpsilon * norm[d];
  }
}

// surfacereflect (smolsurface.c:4821-4869), solution-phase molecule
template <bool kCv>
__device__ __forceinline__ void srf_reflect(const DevPanel& pn, double* pos, const double* crsspt) {
  if (pn.ps == kPsRect) {
    int axis = (int)pn.front[1];
    pos[axis] -= 2.0 * (pos[axis] - crsspt[axis]);
  } else if (pn.ps == kPsDisk || (kCv && pn.ps == kPsTri)) {
    double dot = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) dot += (pos[d] - crsspt[d]) * pn.front[d];
#pragma unroll
    for (int d = 0; d < 3; d++) pos[d] -= 2.0 * pn.front[d] * dot;
  } else {
    double norm[3], dot = 0;
    if (kCv && pn.ps == kPsCyl) srf_linenormal3(pn.point[0], pn.point[1], crsspt, norm);
    else srf_spherenormal(pn.point[0], crsspt, 1, norm);
#pragma unroll
    for (int d = 0; d < 3; d++) dot += (pos[d] - crsspt[d]) * norm[d];
#pragma unroll
    for (int d = 0; d < 3; d++) pos[d] -= 2.0 * norm[d] * dot;
  }
}

// line2nextbox (smolboxes.c:846-892): the next virtual box the segment
// pt1->pt2 enters after box b, or -1
__device__ __forceinline__ int srf_line2nextbox(const Geometry& g, const double* pt1, const double* pt2, int b) {
  if (pos2box(g, pt2[0], pt2[1], pt2[2]) == b) return -1;
  int z1[3];
  z1[2] = b % g.side[2];
  z1[1] = (b / g.side[2]) % g.side[1];
  z1[0] = b / (g.side[2] * g.side[1]);
  const int zb[3] = {z1[0], z1[1], z1[2]};
  double crsmin = 1.01;
  int d2 = 0, boxside2 = 0, flag = 0;
  for (int d = 0; d < 3; d++) {
    if (pt2[d] != pt1[d]) {
      int boxside = (pt2[d] > pt1[d]) ? 1 : 0;
      int sum = zb[d] + boxside;
      if (sum > 0 && sum < g.side[d]) {
        double edge = g.bmin[d] + (double)sum * g.bsize[d];
        double crs = (edge - pt1[d]) / (pt2[d] - pt1[d]);
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
    for (int d = 0; d < 3; d++) {
      if (pt2[d] != pt1[d]) {
        int boxside = (pt2[d] > pt1[d]) ? 1 : 0;
        int sum = zb[d] + boxside;
        if (sum > 0 && sum < g.side[d]) {
          double edge = g.bmin[d] + (double)sum * g.bsize[d];
          double crs = (edge - pt1[d]) / (pt2[d] - pt1[d]);
          if (crs == crsmin && (boxside == 1 || flag == 2)) z1[d] += boxside ? 1 : -1;
        }
      }
    }
  } else {
    if (crsmin == 1.01) return -1;
    z1[d2] += boxside2 ? 1 : -1;
  }
  return (z1[0] * g.side[1] + z1[1]) * g.side[2] + z1[2];
}

// surfaction (smolsurface.c:4330-4384) for a solution-phase molecule at face
// 0/1: returns the action and sets the species after it.  u01 draws the
// uniform of a rate-based action.
template <typename Draw>
__device__ __forceinline__ int srf_action(const DevSurf& sf, int srf, int ident, int face, int* i2, Draw u01,
                                          int* ms2p = nullptr) {
  const DevSrfAct& a = sf.act[(srf * kMaxSpecies + ident) * 2 + face];
  if (a.act != kSaMult) {
    *i2 = a.newspec_fixed;
    if (ms2p) *ms2p = 0;
    return a.act;
  }
  double r = u01();
  int ms2 = 5;
  for (int ms = 0; ms < 6; ms++)
    if (r < a.cumprob[ms]) {
      ms2 = ms;
      break;
    }
  *i2 = a.newspec[ms2];
  if (ms2p) *ms2p = ms2;
  return a.outcome[ms2];
}

// ptinpanel (smolsurface.c:4276-4323) for 3D panels
template <bool kCv>
__device__ __forceinline__ bool srf_ptinpanel(const DevPanel& pn, const double* pt) {
  if (pn.ps == kPsRect) {
    int d = (int)pn.front[2];
    bool in = (pn.point[0][d] <= pt[d] && pt[d] <= pn.point[1][d]) || (pn.point[1][d] <= pt[d] && pt[d] <= pn.point[0][d]);
    d = (d + 1) % 3;
    if (d == (int)pn.front[1]) d = (d + 1) % 3;
    return in && ((pn.point[1][d] <= pt[d] && pt[d] <= pn.point[2][d]) ||
                  (pn.point[2][d] <= pt[d] && pt[d] <= pn.point[1][d]));
  }
  if (pn.ps == kPsSph) return true;
  if (kCv && pn.ps == kPsTri) return srf_ptintriangle(pn, pt);
  if (kCv && pn.ps == kPsCyl) return srf_ptinslab(pn.point[0], pn.point[1], pt);
  if (kCv && pn.ps == kPsHemi) {
    double dot = 0;
    for (int d = 0; d < 3; d++) dot += (pt[d] - pn.point[0][d]) * pn.point[2][d];
    return dot <= 0;
  }
  double len2 = 0, dot = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) len2 += (pt[d] - pn.point[0][d]) * (pt[d] - pn.point[0][d]);
  if (len2 <= pn.point[1][0] * pn.point[1][0]) return true;
#pragma unroll
  for (int d = 0; d < 3; d++) dot += (pt[d] - pn.point[0][d]) * pn.front[d];
  len2 -= dot * dot;
  return len2 <= pn.point[1][0] * pn.point[1][0];
}

// panelnormal (smolsurface.c:4400-4428): outward unit normal of the given face
template <bool kCv>
__device__ __forceinline__ void srf_panelnormal(const DevPanel& pn, const double* pos, int face, double* norm) {
  if (face != kFaceBack) face = kFaceFront;
  if (pn.ps == kPsRect) {
    norm[0] = norm[1] = norm[2] = 0;
    norm[(int)pn.front[1]] =
        ((face == kFaceFront && pn.front[0] == 1) || (face == kFaceBack && pn.front[0] == -1)) ? 1.0 : -1.0;
  } else if (pn.ps == kPsDisk || (kCv && pn.ps == kPsTri)) {
    const double sgn = face == kFaceFront ? 1.0 : -1.0;
#pragma unroll
    for (int d = 0; d < 3; d++) norm[d] = sgn * pn.front[d];
  } else if (kCv && pn.ps == kPsCyl) {
    srf_linenormal3(pn.point[0], pn.point[1], pos, norm);
    if ((face == kFaceBack && pn.front[2] == 1) || (face == kFaceFront && pn.front[2] == -1))
      for (int d = 0; d < 3; d++) norm[d] *= -1;
  } else {
    srf_spherenormal(pn.point[0], pos,
                     ((face == kFaceFront && pn.front[0] == 1) || (face == kFaceBack && pn.front[0] == -1)) ? 1 : -1,
                     norm);
  }
}

// desorbdist (SurfaceParam.c:202-220): rms step times a rational fit of the
// inverse distribution of desorption distances at the uniform y, for
// irreversible desorption or for reversible adsorption
__device__ __forceinline__ double srf_desorbdist(double step, double y, bool revads) {
  const double y2 = y * y;
  double x;
  if (!revads) x = (0.571825 * y - 0.552246 * y2) / (1.0 - 1.53908 * y + 0.546424 * y2);
  else x = (0.729614 * y - 0.70252 * y2) / (1.0 - 1.47494 * y + 0.484371 * y2);
  return step * x;
}

// surfacejump (smolsurface.c:4873-4926) for 3D panels;
// returns 0 when the panel has no jump destination for that face
template <bool kCv>
__device__ __forceinline__ int srf_jump(const DevSurf& sf, const DevPanel& pn, double* pos, double* crsspt,
                                        int face) {
  int p2 = pn.jumpp[face], face2 = pn.jumpf[face];
  if (p2 < 0 || !(face2 == kFaceFront || face2 == kFaceBack)) return 0;
  const DevPanel& q = sf.pnl[p2];
  double delta[3];
  int dir;
  if (pn.ps == kPsRect) {
#pragma unroll
    for (int d = 0; d < 3; d++) delta[d] = 0.5 * (q.point[0][d] + q.point[2][d]) - 0.5 * (pn.point[0][d] + pn.point[2][d]);
    dir = (pn.front[0] == q.front[0]) ? 1 : -1;
  } else if (pn.ps == kPsSph || (kCv && pn.ps == kPsHemi)) {
#pragma unroll
    for (int d = 0; d < 3; d++)
      delta[d] = (crsspt[d] - pn.point[0][d]) * q.point[1][0] / pn.point[1][0] + q.point[0][d] - crsspt[d];
    dir = (pn.front[0] == q.front[0]) ? 1 : -1;
  } else if (kCv && pn.ps == kPsCyl) {
    for (int d = 0; d < 3; d++)
      delta[d] = (crsspt[d] - pn.point[0][d]) * q.point[2][0] / pn.point[2][0] + q.point[0][d] - crsspt[d];
    dir = (pn.front[2] == q.front[2]) ? 1 : -1;
  } else if (kCv && pn.ps == kPsTri) {
    // Geo_TriCenter (Geometry.c:43-53): (1.0/3.0) times the vertex sum
    double dot = 0;
    for (int d = 0; d < 3; d++)
      delta[d] = (1.0 / 3.0) * (q.point[0][d] + q.point[1][d] + q.point[2][d]) -
                 (1.0 / 3.0) * (pn.point[0][d] + pn.point[1][d] + pn.point[2][d]);
    for (int d = 0; d < 3; d++) dot += pn.front[d] * q.front[d];
    dir = dot > 0 ? 1 : -1;
  } else {
    double dot = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) delta[d] = q.point[0][d] - pn.point[0][d];
#pragma unroll
    for (int d = 0; d < 3; d++) dot += pn.front[d] * q.front[d];
    dir = dot > 0 ? 1 : -1;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    crsspt[d] += delta[d];
    pos[d] += delta[d];
  }
  srf_fixpt2panel<kCv>(q, crsspt, face2, 0);
  dir *= (face != face2) ? 1 : -1;
  if (dir == -1) srf_reflect<kCv>(q, pos, crsspt);
  return 1;
}

enum SrfStatus : int {
  kSrfOk = 0, kSrfAbsorbed = 1, kSrfUnsupported = 2, kSrfAdsorbed = 3, kSrfDesorbed = 4, kSrfFlipped = 5
};

// checksurfaces (smolsurface.c:5118-5180) for one solution-phase molecule:
// segment posx -> pos, up to 100 interactions.  ident may change (species
// change on an action); difc gives the motion correction of molchangeident
// (smolmolec.c:2882-2885).  Draws: Philox keyed by (serial, step,
// SMOLRNG_SURFACE, pass<<8 | iteration).  For a desorbed molecule (pass
// SMOLRNG_SURF_DESORBED) this is checksurfaces1mol with crossminimum 0
// (smolsurface.c:5041,5055-5112): crossings at a fraction within VERYCLOSE of
// the segment's start are ignored.
template <bool kCv>
__device__ __forceinline__ int srf_check(const DevSurf& sf, const Geometry& g, const double* difc, double* pos,
                                         const double* posx, int* ident, uint64_t serial, uint32_t step, int pass,
                                         uint32_t key0, uint32_t key1, int* bound = nullptr) {
  double via[3] = {posx[0], posx[1], posx[2]};
  const bool usemin = (pass == SMOLRNG_SURF_DESORBED);
  for (int it = 1;; it++) {
    if (it > 100) {
      pos[0] = posx[0];
      pos[1] = posx[1];
      pos[2] = posx[2];
      return kSrfOk;
    }
    double crossmin = 2, crossmin2 = 2, crssptmin[3] = {0, 0, 0};
    int pmin = -1, facemin = kFaceFront;
    for (int b = pos2box(g, via[0], via[1], via[2]); b >= 0; b = srf_line2nextbox(g, via, pos, b)) {
      for (int k = sf.bpstart[b]; k < sf.bpstart[b + 1]; k++) {
        int p = sf.bplist[k];
        double crsspt[3], cross;
        int face;
        if (srf_linexpanel<kCv>(sf.pnl[p], via, pos, crsspt, &face, nullptr, &cross, nullptr) && cross <= crossmin2 &&
            (!usemin || cross - 0.0 > kVeryClose)) {
          if (cross <= crossmin) {
            crossmin2 = crossmin;
            crossmin = cross;
            pmin = p;
            crssptmin[0] = crsspt[0];
            crssptmin[1] = crsspt[1];
            crssptmin[2] = crsspt[2];
            facemin = face;
          } else {
            crossmin2 = cross;
          }
        }
      }
    }
    if (!(crossmin < 2)) return kSrfOk;
    if (crossmin2 != crossmin && crossmin2 - crossmin < kVeryClose) {
      pos[0] = via[0];
      pos[1] = via[1];
      pos[2] = via[2];
      return kSrfOk;
    }
    // dosurfinteract (smolsurface.c:4930-5051), solution-phase branches
    const DevPanel& pn = sf.pnl[pmin];
    int i = *ident, i2;
    auto draw = [&]() {
      smolrng4 r = smolrng_mol(key0, key1, serial, step, SMOLRNG_SURFACE, ((uint32_t)pass << 8) | (uint32_t)it);
      return smolrng_u01(r.v[0]);
    };
    int ms2;
    int act = srf_action(sf, pn.srf, i, facemin, &i2, draw, &ms2);
    if (act == kSaAdsorb) {
      // adsorption (dosurfinteract, smolsurface.c:5005-5020): onto the panel
      // at the crossing point, on the face of the bound state
      if (!bound || !srf_ptinpanel<kCv>(pn, crssptmin)) return kSrfUnsupported;
#pragma unroll
      for (int d = 0; d < 3; d++) pos[d] = crssptmin[d];
      srf_fixpt2panel<kCv>(pn, pos, ms2 == 1 ? kFaceFront : kFaceBack, sf.epsilon);
      *ident = i2;
      *bound = (pmin << 3) | ms2;
      return kSrfAdsorbed;
    }
    if (act == kSaTrans) {
      srf_fixpt2panel<kCv>(pn, crssptmin, facemin == kFaceFront ? kFaceBack : kFaceFront, sf.epsilon);
    } else if (act == kSaReflect) {
      srf_reflect<kCv>(pn, pos, crssptmin);
      srf_fixpt2panel<kCv>(pn, crssptmin, facemin, sf.epsilon);
    } else if (act == kSaJump) {
      if (!srf_jump<kCv>(sf, pn, pos, crssptmin, facemin))
        srf_fixpt2panel<kCv>(pn, crssptmin, facemin == kFaceFront ? kFaceBack : kFaceFront, sf.epsilon);
    } else if (act == kSaAbsorb) {
      return kSrfAbsorbed;
    } else {
      return kSrfUnsupported;
    }
    if (i2 != i) {  // molchangeident with the crossing point
      if (i2 == 0) return kSrfAbsorbed;
      if (difc[i] > 0) {
        double ratio = sqrt(difc[i2] / difc[i]);
#pragma unroll
        for (int d = 0; d < 3; d++) pos[d] = crssptmin[d] + ratio * (pos[d] - crssptmin[d]);
      }
      *ident = i2;
    }
    via[0] = crssptmin[0];
    via[1] = crssptmin[1];
    via[2] = crssptmin[2];
  }
}

// checksurfacebound (smolsurface.c:5184-5199) for one surface-bound
// molecule: the rate-based action of its state with no face (dosurfinteract,
// smolsurface.c:4930-5051).  On desorption the molecule is placed off the
// panel by the desorption distance along the normal of the face it leaves
// to, checked against the surfaces from the panel (pass DESORBED), and
// becomes a solution-phase molecule: returns kSrfDesorbed with pos and
// ident updated, and *boxp = the box the reference files it in (that of the
// position before the surface check).  bound = (panel << 3) | state.  The
// check may adsorb the new molecule again (dosurfinteract's adsorb branch for
// the resurrected molecule): then *rebound = its (panel << 3) | state, else 0.
// A flip to the other bound state returns kSrfFlipped with *rebound = the
// new (panel << 3) | state and ident updated; absorption returns kSrfAbsorbed.
template <bool kCv>
__device__ __forceinline__ int srf_bound(const DevSurf& sf, const Geometry& g, const double* difc,
                                         const double* difstep, double* pos, int* ident, int bound, uint64_t serial,
                                         uint32_t step, uint32_t key0, uint32_t key1, int* boxp, int* rebound) {
  const int p = bound >> 3, ms = bound & 3;
  const DevPanel& pn = sf.pnl[p];
  const int i = *ident;
  const DevSrfAct& a = sf.actb[(pn.srf * kMaxSpecies + i) * 2 + (ms - 1)];
  if (a.act != kSaMult) return kSrfOk;
  const double r =
      smolrng_u01(smolrng_mol(key0, key1, serial, step, SMOLRNG_SURFACE, (uint32_t)SMOLRNG_SURF_BOUND << 8).v[0]);
  int ms2 = 5;
  for (int m = 0; m < 6; m++)
    if (r < a.cumprob[m]) {
      ms2 = m;
      break;
    }
  const int o = a.outcome[ms2], i2 = a.newspec[ms2];
  if (o == kSaNo) return kSrfOk;
  if (o == kSaFlip) {
    // on-surface flip (smolsurface.c:5044-5048): molchangeident puts the
    // molecule on the face of its new state (smolmolec.c:2898-2901) and
    // dosurfinteract does so again; bound states do not diffuse on this
    // path, so molchangeident makes no motion correction
    if (ms2 != 1 && ms2 != 2) return kSrfUnsupported;
    const int face = ms2 == 1 ? kFaceFront : kFaceBack;
    srf_fixpt2panel<kCv>(pn, pos, face, sf.epsilon);
    srf_fixpt2panel<kCv>(pn, pos, face, sf.epsilon);
    *ident = i2;
    *rebound = (p << 3) | ms2;
    return kSrfFlipped;
  }
  if (o == kSaAbsorb) return kSrfAbsorbed;  // molkill (smolsurface.c:4994-4996)
  if (o != kSaRevdes && o != kSaIrrevdes) return kSrfUnsupported;
  const int face = ms2 == 0 ? kFaceFront : kFaceBack;
  double crsspt[3] = {pos[0], pos[1], pos[2]};
  srf_fixpt2panel<kCv>(pn, crsspt, face, sf.epsilon);
  const double y = smolrng_u01cc(
      smolrng_mol(key0, key1, serial, step, SMOLRNG_SURFACE, (uint32_t)SMOLRNG_SURF_DESORBDIST << 8).v[0]);
  const double x = srf_desorbdist(difstep[i2], y, o == kSaRevdes);
  double norm[3];
  srf_panelnormal<kCv>(pn, crsspt, face, norm);
#pragma unroll
  for (int d = 0; d < 3; d++) pos[d] = crsspt[d] + x * norm[d];
  // the reference files the new molecule in the box of this position, before
  // its surface check (smolsurface.c:5036)
  *boxp = pos2box(g, pos[0], pos[1], pos[2]);
  int i3 = i2, bnd = 0;
  const int st = srf_check<kCv>(sf, g, difc, pos, crsspt, &i3, serial, step, SMOLRNG_SURF_DESORBED, key0, key1, &bnd);
  if (st != kSrfOk && st != kSrfAdsorbed) return kSrfUnsupported;
  *ident = i3;
  *rebound = st == kSrfAdsorbed ? bnd : 0;
  return kSrfDesorbed;
}

// rxnXsurface (smolsurface.c:4388-4424) for two solution-phase molecules: 1
// when a panel between them blocks the reaction.  Rate-based actions draw
// Philox keyed by (serial pair, step, SMOLRNG_RXNSURF, j<<8 | draw index).
template <bool kCv>
__device__ __forceinline__ int srf_rxnblocked(const DevSurf& sf, const Geometry& g, const double* difc, const double* pos1,
                                              const double* pos2, int i1, int i2, uint64_t ser1, uint64_t ser2,
                                              uint32_t step, int j, uint32_t key0, uint32_t key1) {
  double dc1 = difc[i1], dc2 = difc[i2];
  if (dc1 == 0 && dc2 == 0) dc1 = dc2 = 1;
  double rxnpt = dc1 / (dc1 + dc2);
  uint32_t c = 0;
  auto draw = [&]() {
    smolrng4 r = smolrng_pair(key0, key1, ser1, ser2, step, SMOLRNG_RXNSURF, ((uint32_t)j << 8) | c);
    c++;
    return smolrng_u01(r.v[0]);
  };
  auto trans = [&](int srf, int face, int ident) {
    int tmp;
    return srf_action(sf, srf, ident, face, &tmp, draw) == kSaTrans;
  };
  int result = 0;
  for (int b = pos2box(g, pos1[0], pos1[1], pos1[2]); b >= 0 && result == 0; b = srf_line2nextbox(g, pos1, pos2, b)) {
    for (int k = sf.bpstart[b]; k < sf.bpstart[b + 1] && result == 0; k++) {
      const DevPanel& pn = sf.pnl[sf.bplist[k]];
      double crsspt[3], cross, cross2;
      int face1, face2;
      if (srf_linexpanel<kCv>(pn, pos1, pos2, crsspt, &face1, &face2, &cross, &cross2)) {
        if (face1 != face2) {
          if (rxnpt < cross || (rxnpt == cross && face1 == kFaceBack)) result = !trans(pn.srf, face2, i2);
          else result = !trans(pn.srf, face1, i1);
        } else {
          int facein = (face1 == kFaceFront) ? kFaceBack : kFaceFront;
          if (rxnpt < cross || (rxnpt == cross && face1 == kFaceBack))
            result = !(trans(pn.srf, face2, i2) && trans(pn.srf, facein, i2));
          else if (rxnpt < cross2 || (rxnpt == cross2 && face1 == kFaceFront))
            result = !(trans(pn.srf, face1, i1) && trans(pn.srf, face2, i2));
          else result = !(trans(pn.srf, face1, i1) && trans(pn.srf, facein, i1));
        }
      }
    }
  }
  return result;
}

}  // namespace smolamd
#endif
