// Host side of the surfaces (see surfaces_host.h).  Panel geometry follows
// surfaddpanel (smolsurface.c:2432-2640) for 3D rectangles, spheres and
// disks; the box panel lists use the reference's separating-axis tests
// (Geometry.c Geo_RectXaabb3 :1640, Geo_SphsXaabb3 :1736, Geo_DiskXaabb3
// :1918) so each box lists exactly the panels the reference lists, in its
// order (surfaces, then shapes, then panels).
#include "surfaces_host.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

namespace smolamd {

namespace {

enum { MSsoln_ = 0, MSbsoln_ = 5, MSnone_ = 7, MSall_ = 6 };
enum { PFfront_ = 0, PFback_ = 1, PFnone_ = 2, PFboth_ = 3 };
enum { SAreflect_ = 0, SAtrans_ = 1, SAabsorb_ = 2, SAjump_ = 3, SAport_ = 4, SAmult_ = 5, SAno_ = 6,
       SAadsorb_ = 8 };
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

// surfaddpanel for 3D rectangles, spheres and disks; a panel of an existing
// name replaces it
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

int SurfaceSet::build(const Geometry& g, int nspecies, const double* difc_soln, double dt, const int* listof_soln,
                      const bool* mobile_list, int nlist, SurfaceTables& out, char* msg, int msglen) const {
  (void)difc_soln;
  (void)dt;
  (void)listof_soln;
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
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) d.point[a][b] = p.point[a][b];
    for (int b = 0; b < 3; b++) d.front[b] = p.front[b];
    out.pnl.push_back(d);
  }

  // actions of solution-phase molecules at the front and back faces, with
  // the probabilities of rate-based actions (surfupdateparams)
  out.act.assign((size_t)nsrf * kMaxSpecies * 2, DevSrfAct());
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
          if (act == SAabsorb_ || act == SAport_) {
            snprintf(msg, msglen, "surface %s: absorbing and port actions are not implemented on the device path",
                     S.name.c_str());
            return 2;
          }
          continue;
        }
        if (!det) {
          snprintf(msg, msglen, "surface %s: rate-based action without rates", S.name.c_str());
          return 2;
        }
        double prob[kSrfMS1];
        for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
          if (det->datasrc[ms2] == 2) prob[ms2] = det->prob[ms2];
          else if (det->datasrc[ms2] == 1 && det->rate[ms2] != 0) {
            snprintf(msg, msglen,
                     "surface %s: collision probabilities from rates use the reference's adsorption lookup tables, "
                     "which the device path does not restate; give them as probabilities (rate_internal)",
                     S.name.c_str());
            return 3;
          } else prob[ms2] = 0;
        }
        // normalization (smolsurface.c:3758-3772)
        double sum = 0;
        for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
          bool same = (f == PFfront_ && ms2 == MSsoln_) || (f == PFback_ && ms2 == MSbsoln_);
          if (!same) sum += prob[ms2];
        }
        if (sum > 1.0) {
          prob[MSsoln_] = 0;
          for (int ms2 = 0; ms2 < kSrfMS1; ms2++) prob[ms2] /= sum;
        } else {
          prob[f == PFfront_ ? MSsoln_ : MSbsoln_] = 1.0 - sum;
        }
        sum = 0;
        for (int ms2 = 0; ms2 < kSrfMS1; ms2++) {
          sum += prob[ms2];
          a.cumprob[ms2] = sum;
          a.newspec[ms2] = det->newspec[ms2];
          int o;
          if (det->newspec[ms2] == 0) o = SAabsorb_;
          else if (f == PFfront_) o = ms2 == MSsoln_ ? SAreflect_ : ms2 == MSbsoln_ ? SAtrans_ : SAadsorb_;
          else o = ms2 == MSsoln_ ? SAtrans_ : ms2 == MSbsoln_ ? SAreflect_ : SAadsorb_;
          a.outcome[ms2] = o;
          if (prob[ms2] > 0 && (o == SAabsorb_ || o == SAadsorb_)) {
            snprintf(msg, msglen,
                     "surface %s: absorption and adsorption outcomes are not implemented on the device path",
                     S.name.c_str());
            return 2;
          }
        }
      }
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
          else if (ps == 2) in = sphere_x_box(p.point[0], p.point[1][0], v1, v2);
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
