// libsmoldyn-compatible C API over the MI355X particle engine.
//
// Host side of the drop-in boundary: simulation definition calls keep the
// reference's names, argument meaning and ErrorCode behaviour
// (source/Smoldyn/libsmoldyn.cpp), the setup reproduces simupdate
// (source/Smoldyn/smolsim.cpp:2367-2473) for the supported feature set, and
// every time step runs on the device (engine.hip).  There is no CPU fallback:
// a simulation that uses a feature the device path does not implement is
// rejected with ECerror at smolUpdateSim.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <array>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/libsmoldyn_amd.h"
#include "../../include/smoldyn_amd_rng.h"
#include "commands.h"
#include "config_internal.h"
#include "engine.h"
#include "rxnparam.h"
#include "surfaces_host.h"
#include "surrogate_host.h"

using namespace smolamd;

namespace {

constexpr int kMS = 5;  // MSMAX
constexpr int kMS1 = 6; // MSMAX1
constexpr int kStrLen = 512;
constexpr int kGrowCheckSteps = 64;  // smolRunSim checks the lists' fill every this many steps

// last-error state, as the reference keeps it in process globals
// (libsmoldyn.cpp:33-38)
ErrorCode g_errcode = ECok;
ErrorCode g_warncode = ECok;
char g_errfunc[kStrLen] = "";
char g_errstr[kStrLen] = "";
int g_debugmode = 1;

struct Rxn {
  std::string name;
  int order = 0;
  int rct[2] = {0, 0};
  int rctstate[2] = {MSsoln, MSsoln};
  int permit[kMS1 * kMS1] = {0};
  int nprod = 0;
  int prd[kMaxProd] = {0};
  int prdstate[kMaxProd] = {0};
  double rate = -1;
  int multiplicity = -1;
  double bindrad2 = -1, prob = -1, chi = -1;
  int rparamt = RPnone;
  double rparam = 0, unbindrad = -1;
  double prdpos[kMaxProd][3] = {{0}};
  int disable = 0;
  // reaction log (RxnSetLog, smolreact.c:2013-2061): output file name and the
  // serial numbers whose products are logged (logall: every event)
  std::string logfile;
  std::vector<long> logserno;
  bool logall = false;
};

struct PendingMol {
  int ident;
  double pos[3];
  uint64_t serno;
};

}  // namespace

struct smolamd_sim {
  int dim = 3;
  double low[3] = {0, 0, 0}, high[3] = {1, 1, 1};
  char wtype[6] = {'t', 't', 't', 't', 't', 't'};
  double tmin = 0, tmax = 100, dt = 0.01, time = 0, tbreak = 1e300;
  uint32_t stepindex = 0;
  uint64_t key = 0;
  double mpbox = -1, boxsize = -1, accur = 10;
  int maxmol = -1;
  // list capacity of the current engine, and the capacity a rebuild asks for
  // (lists that fill up double, as molexpandlist does)
  long devcap = 0, growcap = 0;
  int rebuilds = 0;                       // engine rebuilds for full lists
  // the reaction events of the step before a rebuild (the old engine's log),
  // served by smolamdGetReactionLog until the next step
  std::vector<LogRow> heldlog;
  bool holdlog = false;
  unsigned long long molsteps_base = 0;   // molecule-steps of engines replaced by a rebuild
  int precision = 64;  // record precision: 64 (FP64 positions) or 32 (FP32 box offsets)
  std::vector<std::string> spname{"empty"};
  std::vector<std::array<double, kMS>> difc{{0, 0, 0, 0, 0}};
  std::vector<std::array<int, kMS>> listlookup{{-1, -1, -1, -1, -1}};
  std::vector<std::string> listname;
  std::vector<PendingMol> pending;
  uint64_t serno = 1;
  std::vector<Rxn> rxn[3];
  std::vector<std::vector<int>> table[3];  // [order][ident pack] -> reaction indices
  bool ok = false;
  bool started = false;
  DeviceState* dev = nullptr;
  std::vector<int> diffuselist;
  Geometry geo;
  DevTables tab;
  std::vector<double> gtable;
  int counts[kMaxSpecies] = {0};
  SurfaceSet surf;
  cmd::Commands cmds;  // run-time commands, output files and data tables
  int obsrows = 0;     // observation rows queued on the device since the last flush
  // device inputs of the last update, kept for ensembles
  struct Prepared {
    EngineConfig cfg;
    DevTables tab;
    Geometry geo;
    std::vector<std::vector<int>> ident;
    std::vector<std::vector<double>> pos;
    std::vector<std::vector<uint64_t>> ser;
  } prep;
};

namespace {

void set_error(const char* func, ErrorCode code, const char* fmt, ...) {
  char msg[kStrLen];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  smolSetError(func, code, msg, "");
}

// smolSetErrorNT (libsmoldyn.cpp:111-122): the error recorded without the
// debug report; returns the code
int set_error_nt(const char* func, ErrorCode code, const char* msg) {
  if (code != ECsame) {
    g_errcode = code;
    g_warncode = code >= ECwarning ? code : ECok;
    snprintf(g_errstr, kStrLen, "%s", msg ? msg : "");
  }
  snprintf(g_errfunc, kStrLen, "%s", func ? func : "");
  return (int)code;
}

// LCHECK(..., ECsame, NULL) of the reference: the error an inner lookup
// recorded stays (code and message), now named for the calling function
ErrorCode same_error(const char* func) {
  smolSetError(func, ECsame, nullptr, "");
  return g_errcode;
}

#define LCHECK(cond, func, code, ...)     \
  do {                                    \
    if (!(cond)) {                        \
      set_error(func, code, __VA_ARGS__); \
      if (code < ECwarning) return code;  \
    }                                     \
  } while (0)

int species_index(simptr sim, const char* name) {
  if (!name) return -1;
  for (size_t i = 1; i < sim->spname.size(); i++)
    if (sim->spname[i] == name) return (int)i;
  return -1;
}

int list_index(simptr sim, const char* name) {
  for (size_t l = 0; l < sim->listname.size(); l++)
    if (sim->listname[l] == name) return (int)l;
  return -1;
}

void invalidate(simptr sim) { sim->ok = false; }

bool ismobile(simptr sim, int i, int ms) { return sim->difc[i][ms] > 0; }

// RxnSetPermit (smolreact.c:1899-1922)
void set_permit(Rxn& r) {
  int nms = r.order == 1 ? kMS1 : kMS1 * kMS1;
  for (int ms = 0; ms < nms; ms++) {
    int msl[2];
    if (r.order == 1) msl[0] = ms;
    else {
      msl[0] = ms / kMS1;
      msl[1] = ms % kMS1;
    }
    bool set = true;
    for (int o = 0; o < r.order && set; o++)
      if (!(r.rctstate[o] == MSall || r.rctstate[o] == msl[o])) set = false;
    if (set) r.permit[ms] = 1;
  }
  if (r.order == 2 && r.rct[0] == r.rct[1] && r.rctstate[0] != r.rctstate[1]) {
    for (int ms = 0; ms < nms; ms++) {
      int a = ms / kMS1, b = ms % kMS1;
      if ((r.rctstate[1] == MSall || r.rctstate[1] == a) && (r.rctstate[0] == MSall || r.rctstate[0] == b))
        r.permit[ms] = 1;
    }
  }
}

int pack(int order, const int* id) { return order == 1 ? id[0] : id[0] * kMaxSpecies + id[1]; }

// findreverserxn (smolreact.c:345-381)
int find_reverse(simptr sim, int order, int r, int* optr, int* rptr) {
  Rxn& rx = sim->rxn[order][r];
  int orderr = 0, rret = 0, rev = 0;
  if (rx.nprod == 0 || rx.nprod >= 3 || sim->rxn[rx.nprod].empty()) {
    rev = 0;
  } else {
    orderr = rx.nprod;
    int identr = pack(orderr, rx.prd);
    int mstater = orderr == 1 ? rx.prdstate[0] : rx.prdstate[0] * kMS1 + rx.prdstate[1];
    const std::vector<int>& tr = sim->table[orderr][identr];
    for (size_t j = 0; j < tr.size(); j++) {
      int rr = tr[j];
      Rxn& rxr = sim->rxn[orderr][rr];
      if (!rxr.permit[mstater]) continue;
      if (rev != 1 && rxr.nprod == order) {
        bool same = order == 1 ? rx.rct[0] == rxr.prd[0]
                               : ((rx.rct[0] == rxr.prd[0] && rx.rct[1] == rxr.prd[1]) ||
                                  (rx.rct[0] == rxr.prd[1] && rx.rct[1] == rxr.prd[0]));
        if (same) {
          int identrprd = pack(order, rxr.prd);
          int mstaterprd = order == 1 ? rxr.prdstate[0] : rxr.prdstate[0] * kMS1 + rxr.prdstate[1];
          const std::vector<int>& tf = sim->table[order][identrprd];
          for (size_t jr = 0; jr < tf.size(); jr++)
            if (tf[jr] == r && sim->rxn[order][r].permit[mstaterprd]) {
              rev = 1;
              rret = rr;
            }
        }
      }
      if (!rev) {
        rev = 2;
        rret = rr;
      }
    }
  }
  *optr = orderr;
  *rptr = rret;
  return rev;
}

double difcsum(simptr sim, int i1, int ms1, int i2, int ms2) {
  double sum = 0;
  if (i1) sum += sim->difc[i1][ms1 >= kMS ? MSsoln : ms1];
  if (i2) sum += sim->difc[i2][ms2 >= kMS ? MSsoln : ms2];
  return sum;
}

// rxnreactantstate, order 2, converting bsoln to soln (smolreact.c:277-343)
bool rctstate2(const Rxn& r, int* ms1, int* ms2) {
  const int* p = r.permit;
  if (p[MSsoln * kMS1 + MSsoln] || p[MSsoln * kMS1 + MSbsoln] || p[MSbsoln * kMS1 + MSsoln] ||
      p[MSbsoln * kMS1 + MSbsoln]) {
    *ms1 = *ms2 = MSsoln;
    return true;
  }
  for (int a = 0; a < kMS1; a++)
    if (p[a * kMS1 + MSsoln]) {
      *ms1 = a == MSbsoln ? MSsoln : a;
      *ms2 = MSsoln;
      return true;
    }
  for (int b = 0; b < kMS1; b++)
    if (p[MSsoln * kMS1 + b]) {
      *ms1 = MSsoln;
      *ms2 = b == MSbsoln ? MSsoln : b;
      return true;
    }
  for (int mp = 0; mp < kMS1 * kMS1; mp++)
    if (p[mp]) {
      *ms1 = mp / kMS1;
      *ms2 = mp % kMS1;
      if (*ms1 == MSbsoln) *ms1 = MSsoln;
      if (*ms2 == MSbsoln) *ms2 = MSsoln;
      return true;
    }
  *ms1 = *ms2 = MSnone;
  return false;
}

// rxnsetrate (smolreact.c:1306-1412)
int set_rate(simptr sim, int order, int r) {
  Rxn& rx = sim->rxn[order][r];
  int o2, r2, rev;
  if (rx.disable) {
    rx.prob = 0;
  } else if (rx.rparamt == RPconfspread) {
    if (rx.rate < 0) return 1;
    rx.prob = 1.0 - exp(-sim->dt * rx.rate * rx.multiplicity);
  } else if (order == 0) {
    // events per step in the system volume (systemvolume, smolwall.c:50-57)
    if (rx.rate < 0) return 1;
    double vol = 1.0;
    for (int d = 0; d < sim->dim; d++) vol *= sim->high[d] - sim->low[d];
    rx.prob = rx.rate * rx.multiplicity * sim->dt * vol;
  } else if (order == 1) {
    int ms;
    for (ms = 0; ms < kMS && !rx.permit[ms]; ms++) {
    }
    if (rx.rate < 0) return 1;
    if (rx.rate == 0 || ms == kMS) {
      rx.prob = 0;
    } else {
      const std::vector<int>& t = sim->table[1][rx.rct[0]];
      double sum = 0;
      for (int rr : t) {
        Rxn& x2 = sim->rxn[1][rr];
        if (x2.permit[ms] && x2.rate > 0) sum += x2.rate * x2.multiplicity;
      }
      rx.prob = rx.rate * rx.multiplicity / sum * (1.0 - exp(-sum * sim->dt));
      double product = 1.0;
      for (int rr : t) {
        Rxn& x2 = sim->rxn[1][rr];
        if (&x2 == &rx) break;
        if (x2.permit[ms]) {
          double prob = x2.prob * product;
          product *= 1.0 - prob;
        }
      }
      rx.prob /= product;
      if (!(rx.prob >= 0 && rx.prob <= 1)) return 5;
    }
    rev = find_reverse(sim, 1, r, &o2, &r2);
    if (rev > 0 && o2 == 2 && rx.rparamt == RPnone) {
      rx.rparamt = RPpgemmaxw;
      rx.rparam = 0.2;
    }
  } else {
    if (rx.rate < 0) {
      if (rx.prob < 0) rx.prob = 1;
      return 1;
    }
    int ms1, ms2;
    bool permit = rctstate2(rx, &ms1, &ms2);
    double rate3 = rx.rate * rx.multiplicity;
    if (rx.rct[0] == rx.rct[1]) rate3 *= 2;
    if ((ms1 == MSsoln && ms2 != MSsoln) || (ms1 != MSsoln && ms2 == MSsoln)) rate3 *= 2;
    double dsum = difcsum(sim, rx.rct[0], ms1, rx.rct[1], ms2);
    rev = find_reverse(sim, 2, r, &o2, &r2);
    if (rev > 0 && o2 == 2 && rx.rparamt == RPnone) {
      rx.rparamt = RPpgemmaxw;
      rx.rparam = 0.2;
    }
    int rparamt = RPnone;
    double rparam = 0;
    if (rev == 1) {
      Rxn& rv = sim->rxn[o2][r2];
      if (rv.rparamt == RPnone) {
        rv.rparamt = RPpgemmaxw;
        rv.rparam = 0.2;
      }
      rparamt = rv.rparamt;
      rparam = rv.rparam;
    }
    if (rx.prob < 0) rx.prob = 1;
    double dtt = sim->dt;
    if (!permit) rx.bindrad2 = 0;
    else if (rate3 <= 0) rx.bindrad2 = 0;
    else if (dsum <= 0) return 4;
    else if (rparamt == RPunbindrad)
      rx.bindrad2 = rxnparam::bindingradiusprob(rate3, dtt, dsum, rparam, 0, rx.chi, &rx.prob);
    else if (rparamt == RPratio && surrogate::enabled() && rx.chi < 0 && rx.prob == 1)
      rx.bindrad2 = surrogate::binding_radius(rate3, dtt, dsum, rparam, nullptr);
    else if (rparamt == RPratio)
      rx.bindrad2 = rxnparam::bindingradiusprob(rate3, dtt, dsum, rparam, 1, rx.chi, &rx.prob);
    else if (rparamt == RPpgem)
      rx.bindrad2 = rxnparam::bindingradiusprob(rate3 * (1.0 - rparam), dtt, dsum, -1, 0, rx.chi, &rx.prob);
    else if (rparamt == RPpgemmax || rparamt == RPpgemmaxw) {
      rx.bindrad2 = rxnparam::bindingradiusprob(rate3, dtt, dsum, 0, 0, rx.chi, &rx.prob);
      double ub = rxnparam::unbindingradius(rparam, dtt, dsum, rx.bindrad2);
      if (ub > 0) rx.bindrad2 = rxnparam::bindingradiusprob(rate3 * (1.0 - rparam), dtt, dsum, -1, 0, rx.chi, &rx.prob);
    } else
      rx.bindrad2 = rxnparam::bindingradiusprob(rate3, dtt, dsum, -1, 0, rx.chi, &rx.prob);
    rx.bindrad2 *= rx.bindrad2;
  }
  return 0;
}

// rxnsetproduct (smolreact.c:1431-1575), except bounce placement
int set_product(simptr sim, int order, int r) {
  Rxn& rx = sim->rxn[order][r];
  int nprod = rx.nprod, rparamt = rx.rparamt, dim = sim->dim;
  double rpar = rx.rparam;
  if (nprod == 0) {
    if (rparamt == RPnone || rparamt == RPirrev || rparamt == RPconfspread) rx.unbindrad = 0;
    else return 1;
  } else if (rparamt == RPoffset || rparamt == RPfixed) {
    // the offsets are the caller's; unbinding radius as rxnsetproduct computes it (smolreact.c:1451-1458)
    double dist = 0;
    if (nprod == 1)
      for (int d = 0; d < dim; d++) dist = rx.prdpos[0][d];
    else
      for (int d = 0; d < dim; d++) dist += rx.prdpos[0][d] - rx.prdpos[1][d];
    rx.unbindrad = sqrt(dist);
  } else if (nprod == 1) {
    if (rparamt == RPnone || rparamt == RPirrev || rparamt == RPconfspread) {
      rx.unbindrad = 0;
      for (int d = 0; d < dim; d++) rx.prdpos[0][d] = 0;
    } else
      return 2;
  } else if (rparamt == RPbounce) {
    // nprod >= 2, bounce (smolreact.c:1470-1494); the reflection method
    // (rparam -2) keeps the rparam < 0 offsets for its fallback
    int ms1 = rx.rctstate[0], ms2 = rx.rctstate[1];
    if (ms1 == MSbsoln || ms1 == MSnone || ms1 == MSall || ms1 == MSsome) ms1 = MSsoln;
    if (ms2 == MSbsoln || ms2 == MSnone || ms2 == MSall || ms2 == MSsome) ms2 = MSsoln;
    double dc1 = difcsum(sim, rx.rct[0], ms1, 0, MSnone), dc2 = difcsum(sim, rx.rct[1], ms2, 0, MSnone);
    const double dsum = dc1 + dc2;
    if (dsum == 0) dc1 = dc2 = 0.5;
    else {
      dc1 /= dsum;
      dc2 /= dsum;
    }
    const double len = rpar >= 0 ? rpar : sqrt(rx.bindrad2);
    rx.unbindrad = rpar;
    rx.prdpos[0][0] = len * dc1;
    rx.prdpos[1][0] = len * dc2;
    for (int p = 2; p < nprod; p++) rx.prdpos[p][0] = 0;
  } else {
    int ms1 = rx.prdstate[0] == MSbsoln ? MSsoln : rx.prdstate[0];
    int ms2 = rx.prdstate[1] == MSbsoln ? MSsoln : rx.prdstate[1];
    double dc1 = difcsum(sim, rx.prd[0], ms1, 0, MSnone), dc2 = difcsum(sim, rx.prd[1], ms2, 0, MSnone);
    double dsum = dc1 + dc2;
    if (dsum == 0) dc1 = dc2 = 0.5;
    else {
      dc1 /= dsum;
      dc2 /= dsum;
    }
    int orderr, rr;
    int rev = find_reverse(sim, order, r, &orderr, &rr);
    if (rparamt == RPirrev || rparamt == RPconfspread) {
      rx.unbindrad = 0;
      for (int p = 0; p < nprod; p++)
        for (int d = 0; d < dim; d++) rx.prdpos[p][d] = 0;
    } else if (rev == 0) {
      if (rparamt == RPpgem || rparamt == RPpgemmax || rparamt == RPpgemmaxw || rparamt == RPratio ||
          rparamt == RPpgem2 || rparamt == RPpgemmax2 || rparamt == RPratio2)
        return 3;
      if (rparamt == RPunbindrad) {
        rx.unbindrad = rpar;
        rx.prdpos[0][0] = rpar * dc1;
        rx.prdpos[1][0] = -rpar * dc2;
      } else {
        rx.unbindrad = 0;
        rx.prdpos[0][0] = 0;
        rx.prdpos[1][0] = 0;
      }
    } else {
      Rxn& rxr = sim->rxn[orderr][rr];
      double bindradr = rxr.bindrad2 >= 0 ? sqrt(rxr.bindrad2) : -1;
      if (rparamt == RPnone) return 5;
      if (rparamt == RPunbindrad) {
        rx.unbindrad = rpar;
        rx.prdpos[0][0] = rpar * dc1;
        rx.prdpos[1][0] = -rpar * dc2;
      } else if (rxr.bindrad2 < 0) {
        return 6;
      } else if (rxr.bindrad2 == 0) {
        rx.unbindrad = 0;
        rx.prdpos[0][0] = 0;
        rx.prdpos[1][0] = 0;
      } else if (rparamt == RPratio || rparamt == RPratio2) {
        rx.unbindrad = rpar * bindradr;
        rx.prdpos[0][0] = rpar * bindradr * dc1;
        rx.prdpos[1][0] = -rpar * bindradr * dc2;
      } else if (dsum <= 0) {
        return 4;
      } else if (rparamt == RPpgem || rparamt == RPpgem2) {
        rpar = rxnparam::unbindingradius(rpar, sim->dt, dsum, bindradr);
        if (rpar < 0) return 7;
        rx.unbindrad = rpar;
        rx.prdpos[0][0] = rpar * dc1;
        rx.prdpos[1][0] = -rpar * dc2;
      } else if (rparamt == RPpgemmax || rparamt == RPpgemmaxw || rparamt == RPpgemmax2) {
        rpar = rxnparam::unbindingradius(rpar, sim->dt, dsum, bindradr);
        if (rpar == -2) return 9;
        if (rpar <= 0) rx.unbindrad = 0;
        else {
          rx.unbindrad = rpar;
          rx.prdpos[0][0] = rpar * dc1;
          rx.prdpos[1][0] = -rpar * dc2;
        }
      } else
        return 10;
    }
  }
  return 0;
}

// Gaussian quantile table (random2.c:339-353 with math2.c's erfn/inversefn):
// the same construction as the reference, unshuffled because every index is
// drawn from the counter stream.
// the reference's gammalnD returns this constant for x == 0.5 (math2.c:268)
constexpr double kLnGammaHalf = 0.572364942;

double gammp_half(double x) {
  const double a = 0.5;
  double sum, del, ap, eps = 3e-7;
  double gold = 0, g = 1, fac = 1, b1 = 1, b0 = 0, anf, ana, an, a1, a0 = 1;
  if (x < 0) return -1;
  if (x == 0) return 0;
  if (x < a + 1) {
    ap = a;
    del = sum = 1.0 / a;
    while (fabs(del) > fabs(sum) * eps && ap - a < 100) {
      ap += 1;
      del *= x / ap;
      sum += del;
    }
    return sum * exp(-x + a * log(x) - kLnGammaHalf);
  }
  a1 = x;
  for (an = 1; an < 100; an++) {
    ana = an - a;
    a0 = (a1 + a0 * ana) * fac;
    b0 = (b1 + b0 * ana) * fac;
    anf = an * fac;
    a1 = x * a0 + anf * a1;
    b1 = x * b0 + anf * b1;
    if (a1) {
      fac = 1.0 / a1;
      g = b1 * fac;
      if (fabs((g - gold) / g) < eps) return 1.0 - exp(-x + a * log(x) - kLnGammaHalf) * g;
      gold = g;
    }
  }
  return -1;
}

double erfn(double x) { return x < 0 ? -gammp_half(x * x) : gammp_half(x * x); }

void gauss_table(std::vector<double>& a, int n) {
  a.assign(n, 0.0);
  double dy = 2.0 / n;
  for (int i = 0; i < n / 2; i++) {
    double y = (i + 0.5) * dy - 1.0;
    double x1 = -20, x2 = 20, dx, y2;
    if (erfn(x1) < erfn(x2)) dx = x2 - x1;
    else {
      dx = x1 - x2;
      x1 = x2;
    }
    for (int k = 30; k > 0; k--) {
      dx *= 0.5;
      x2 = x1 + dx;
      y2 = erfn(x2);
      if (y2 < y) x1 = x2;
    }
    a[i] = 1.41421356237 * (x1 + 0.5 * dx);
  }
  for (int i = n / 2; i < n; i++) a[i] = -a[n - i - 1];
}

// Setup of simupdate (smolsim.cpp:2367-2473) for the device path.  With
// for_batch the device inputs are kept in sim->prep and no engine is
// created (the simulation becomes one member of an ensemble).
// rebuild: a new engine for a running simulation (grow_lists): the box grid
// is kept, the molecules are the current ones in sim->pending
ErrorCode do_update(simptr sim, const char* fn, bool for_batch = false, bool rebuild = false) {
  if (sim->ok && !for_batch && !rebuild) return ECok;
  LCHECK(!sim->started || rebuild, fn, ECerror, "the simulation definition changed after stepping started");
  LCHECK(sim->dim == 3, fn, ECerror, "the device engine runs 3-D systems");
  int nsp = (int)sim->spname.size();
  LCHECK(nsp <= kMaxSpecies, fn, ECbounds, "too many species for the device tables (%d)", kMaxSpecies);
  // molsupdatelists (smolmolec.c:2587-2651)
  std::vector<std::array<int, kMS>> lookup = sim->listlookup;
  std::vector<std::string> lists = sim->listname;
  if (lists.empty() && !sim->pending.empty() && nsp > 1) {
    bool ndif = false, nfix = false;
    for (int i = 1; i < nsp; i++)
      for (int ms = 0; ms < kMS; ms++) (ismobile(sim, i, ms) ? ndif : nfix) = true;
    if (ndif) {
      lists.push_back("diffuselist");
      for (int i = 1; i < nsp; i++)
        for (int ms = 0; ms < kMS; ms++)
          if (ismobile(sim, i, ms)) lookup[i][ms] = (int)lists.size() - 1;
    }
    if (nfix) {
      lists.push_back("fixedlist");
      for (int i = 1; i < nsp; i++)
        for (int ms = 0; ms < kMS; ms++)
          if (!ismobile(sim, i, ms)) lookup[i][ms] = (int)lists.size() - 1;
    }
  }
  bool unset = false;
  for (int i = 1; i < nsp; i++)
    for (int ms = 0; ms < kMS; ms++)
      if (lookup[i][ms] < 0) unset = true;
  if (unset) {
    int ull = -1;
    for (size_t l = 0; l < lists.size(); l++)
      if (lists[l] == "unassignedlist") ull = (int)l;
    if (ull < 0) {
      lists.push_back("unassignedlist");
      ull = (int)lists.size() - 1;
    }
    for (int i = 1; i < nsp; i++)
      for (int ms = 0; ms < kMS; ms++)
        if (lookup[i][ms] < 0) lookup[i][ms] = ull;
  }
  int nlist = (int)lists.size();
  LCHECK(nlist >= 1 && nlist <= kMaxList, fn, ECbounds, "molecule list count %d outside 1..%d", nlist, kMaxList);
  sim->listname = lists;
  sim->listlookup = lookup;
  sim->diffuselist.assign(nlist, 0);
  for (int i = 0; i < nsp; i++)
    for (int ms = 0; ms < kMS; ms++)
      if (ismobile(sim, i, ms) && lookup[i][ms] >= 0) sim->diffuselist[lookup[i][ms]] = 1;
  gauss_table(sim->gtable, 4096);

  // boxesupdatelists (smolboxes.c:717-819)
  Geometry& g = sim->geo;
  const Geometry kept = g;  // a rebuild keeps the virtual boxes of the first update
  memset(&g, 0, sizeof(g));
  double mpbox = sim->mpbox;
  if (mpbox < 0 && sim->boxsize < 0) mpbox = 4;
  if (mpbox < 0) mpbox = 0;
  if (mpbox <= 0 && sim->boxsize <= 0) mpbox = 5;
  double flt1;
  if (mpbox > 0) {
    double vol = 1.0;
    for (int d = 0; d < 3; d++) vol *= sim->high[d] - sim->low[d];
    flt1 = pow((double)sim->pending.size() / mpbox / vol, 1.0 / 3);
  } else {
    flt1 = 1.0 / sim->boxsize;
  }
  long nbox = 1;
  for (int d = 0; d < 3; d++) {
    g.low[d] = sim->low[d];
    g.high[d] = sim->high[d];
    g.width[d] = sim->high[d] - sim->low[d];
    g.bmin[d] = sim->low[d];
    g.side[d] = (int)ceil((sim->high[d] - sim->low[d]) * flt1);
    if (!g.side[d]) g.side[d] = 1;
    if (rebuild) g.side[d] = kept.side[d];
    g.bsize[d] = (sim->high[d] - sim->low[d]) / g.side[d];
    nbox *= g.side[d];
  }
  LCHECK(nbox < (1L << 26), fn, ECbounds, "virtual box grid too large (%ld boxes)", nbox);
  g.nbox = (int)nbox;
  for (int w = 0; w < 6; w++)
    g.wall[w] = sim->wtype[w] == 'r'   ? kWallReflect
                : sim->wtype[w] == 'p' ? kWallPeriodic
                : sim->wtype[w] == 'a' ? kWallAbsorb
                                       : kWallTransparent;
  // neighborZV type 7 reads the first dim entries of the per-wall periodic
  // flags as per-dimension flags (Zn.c:263, smolboxes.c:772-776)
  LCHECK(sim->accur >= 9, fn, ECerror, "accuracy below 9 selects neighbor lists the device path does not build");
  // no neighbor wrapping once surfaces exist (smolboxes.c:772-776)
  for (int d = 0; d < 3; d++) g.wrapflag[d] = (sim->wtype[d] == 'p' && !sim->surf.enabled());

  // reactions (rxnsupdate, smolreact.c:2384-2401)
  for (int order = 0; order <= 2; order++)
    for (size_t r = 0; r < sim->rxn[order].size(); r++) {
      int er = set_rate(sim, order, (int)r);
      LCHECK(er <= 1, fn, ECerror, "error setting rate for reaction %s (code %d)", sim->rxn[order][r].name.c_str(), er);
    }
  LCHECK(sim->rxn[0].size() <= (size_t)kMaxRxn0, fn, ECbounds, "too many zeroth-order reactions");
  for (auto& r0 : sim->rxn[0])
    LCHECK(r0.nprod <= 1, fn, ECerror,
           "zeroth-order reaction %s: more than one product is not implemented on the device path",
           r0.name.c_str());
  for (int order = 0; order <= 2; order++)
    for (size_t r = 0; r < sim->rxn[order].size(); r++) {
      int er = set_product(sim, order, (int)r);
      LCHECK(er == 0, fn, ECerror, "product placement of reaction %s is not supported here (code %d)",
             sim->rxn[order][r].name.c_str(), er);
    }

  // device tables
  DevTables& t = sim->tab;
  memset(&t, 0, sizeof(t));
  t.nspecies = nsp;
  t.nlist = nlist;
  t.dim = 3;
  t.ngauss = 4096;
  for (int i = 0; i < nsp; i++) {
    t.difc[i] = sim->difc[i][MSsoln];
    t.difstep[i] = sqrt(2.0 * sim->difc[i][MSsoln] * sim->dt);
    t.listof[i] = i ? lookup[i][MSsoln] : 0;
  }
  for (int l = 0; l < nlist; l++) t.diffuselist[l] = sim->diffuselist[l];
  for (int i = 1; i < nsp; i++) {
    const std::vector<int>& tb = sim->table[1].size() > (size_t)i ? sim->table[1][i] : std::vector<int>();
    LCHECK((int)tb.size() <= kMaxRxnPerSet, fn, ECbounds, "too many first-order reactions for species %d", i);
    t.nuni[i] = (int)tb.size();
    for (size_t j = 0; j < tb.size(); j++) t.uni[i][j] = tb[j];
  }
  for (int i1 = 1; i1 < nsp; i1++)
    for (int i2 = 1; i2 < nsp; i2++) {
      int k = i1 * kMaxSpecies + i2;
      const std::vector<int>& tb = sim->table[2].size() > (size_t)k ? sim->table[2][k] : std::vector<int>();
      LCHECK((int)tb.size() <= kMaxRxnPerSet, fn, ECbounds, "too many reactions for a species pair");
      t.nbi[k] = (int)tb.size();
      for (size_t j = 0; j < tb.size(); j++) t.bi[k][j] = tb[j];
    }
  LCHECK(sim->rxn[1].size() <= (size_t)kMaxRxn && sim->rxn[2].size() <= (size_t)kMaxRxn, fn, ECbounds,
         "too many reactions");
  t.nrxn0 = (int)sim->rxn[0].size();
  for (int d = 0; d < 3; d++) {
    t.zlo[d] = sim->low[d];
    t.zhi[d] = sim->high[d];
  }
  for (int r = 0; r < t.nrxn0; r++) {
    const Rxn& rx = sim->rxn[0][(size_t)r];
    DevRxn& dr = t.rxn0[r];
    dr.order = 0;
    dr.nprod = rx.nprod;
    for (int p = 0; p < rx.nprod; p++) {
      dr.prd[p] = rx.prd[p];
      LCHECK(rx.prdstate[p] == MSsoln || rx.prdstate[p] == MSbsoln, fn, ECerror,
             "surface-bound products need the surface path");
      dr.prdlist[p] = lookup[rx.prd[p]][MSsoln];
      LCHECK(rx.prdpos[p][0] == 0 && rx.prdpos[p][1] == 0 && rx.prdpos[p][2] == 0, fn, ECerror,
             "zeroth-order reaction %s: product offsets are not implemented on the device path", rx.name.c_str());
    }
    const double lam = rx.prob > 0 ? rx.prob : 0;
    t.zm[r] = lam > 0 ? (int)ceil(lam / SMOLRNG_ZERO_CHUNK) : 0;
    t.zg[r] = t.zm[r] ? exp(-lam / t.zm[r]) : 1.0;
  }
  for (int order = 1; order <= 2; order++)
    for (size_t r = 0; r < sim->rxn[order].size(); r++) {
      Rxn& rx = sim->rxn[order][r];
      DevRxn& dr = order == 1 ? t.rxn1[r] : t.rxn2[r];
      dr.order = order;
      dr.rct0 = rx.rct[0];
      dr.rct1 = order == 2 ? rx.rct[1] : 0;
      dr.nprod = rx.nprod;
      dr.bindrad2 = rx.bindrad2;
      dr.prob = rx.prob;
      dr.confspread = rx.rparamt == RPconfspread;
      dr.fixed = rx.rparamt == RPfixed;
      dr.bounce = order == 2 && rx.rparamt == RPbounce;
      dr.bparam = rx.rparam;
      dr.permit_soln = order == 1 ? rx.permit[MSsoln] : rx.permit[MSsoln * kMS1 + MSsoln];
      for (int p = 0; p < rx.nprod; p++) {
        dr.prd[p] = rx.prd[p];
        LCHECK(rx.prdstate[p] == MSsoln || rx.prdstate[p] == MSbsoln, fn, ECerror,
               "surface-bound products need the surface path");
        dr.prdlist[p] = lookup[rx.prd[p]][MSsoln];
        dr.prdpos[p] = rx.prdpos[p][0];
        for (int d = 0; d < 3; d++) dr.prdvec[p][d] = rx.prdpos[p][d];
        if (dr.fixed || dr.bounce) continue;
        if (rx.prdpos[p][0] != 0 || rx.prdpos[p][1] != 0 || rx.prdpos[p][2] != 0) dr.rotate = 1;
        if (rx.prdpos[p][1] != 0 || rx.prdpos[p][2] != 0) dr.offrot = 1;
      }
    }
  // pair-search screening tables: which partner species react, and the
  // largest binding radius of each species
  for (int i1 = 1; i1 < nsp; i1++) {
    t.reactmask[i1] = 0;
    t.maxbind2[i1] = 0;
    for (int i2 = 1; i2 < nsp; i2++) {
      int k = i1 * kMaxSpecies + i2;
      for (int j = 0; j < t.nbi[k]; j++) {
        const DevRxn& dr = t.rxn2[t.bi[k][j]];
        if (dr.bindrad2 > 0 && dr.prob > 0 && dr.permit_soln) {
          t.reactmask[i1] |= 1u << i2;
          if (dr.bindrad2 > t.maxbind2[i1]) t.maxbind2[i1] = dr.bindrad2;
        }
      }
    }
  }
  {
    double ext = 0;
    for (int d = 0; d < 3; d++) {
      ext = fmax(ext, fabs(sim->low[d]));
      ext = fmax(ext, fabs(sim->high[d]));
      ext = fmax(ext, sim->high[d] - sim->low[d]);
    }
    // Molecules with a coordinate beyond 2*ext bypass the FP32 screen (the
    // engine flags them).  For the others, FP32 rounding of the coordinates
    // (<= 2*ext), of the periodic shift (<= ext) and of their difference
    // perturbs a distance by less than 6 * ext * 2^-23; the margin is 32 * ext
    // * 2^-23.
    g.prefilter_margin = 32.0 * ext * ldexp(1.0, -23);
  }
  // rxnsupdatelists order 2 (smolreact.c:2352-2377) and the (ll1,ll2) loop rank
  for (size_t r = 0; r < sim->rxn[2].size(); r++) {
    Rxn& rx = sim->rxn[2][r];
    for (int ms1 = 0; ms1 < kMS1; ms1++)
      for (int ms2 = 0; ms2 < kMS1; ms2++)
        if (rx.permit[ms1 * kMS1 + ms2] && rx.prob != 0 && (rx.rate > 0 || rx.bindrad2 > 0)) {
          int l1 = lookup[rx.rct[0]][ms1 == MSbsoln ? MSsoln : ms1];
          int l2 = lookup[rx.rct[1]][ms2 == MSbsoln ? MSsoln : ms2];
          t.rxnmol2[l1 * kMaxList + l2] = 1;
          t.rxnmol2[l2 * kMaxList + l1] = 1;
        }
  }
  {
    int rank = 0;
    for (int l1 = 0; l1 < nlist; l1++)
      for (int l2 = l1; l2 < nlist; l2++) t.pairindex[l1 * kMaxList + l2] = rank++;
    LCHECK(rank <= 16, fn, ECbounds, "too many molecule lists for the pair key");
    for (int l1 = 0; l1 < nlist; l1++)
      for (int l2 = l1; l2 < nlist; l2++)
        LCHECK(!t.rxnmol2[l1 * kMaxList + l2] || (t.diffuselist[l1] && t.diffuselist[l2]), fn, ECerror,
               "bimolecular reactions with an immobile list are not implemented on the device path");
  }
  // initial live order: molsort appends the molecules in the order they sit
  // in the dead list's resurrected region, i.e. newest first
  std::vector<std::vector<int>> ident(nlist);
  std::vector<std::vector<double>> pos(nlist);
  std::vector<std::vector<uint64_t>> ser(nlist);
  memset(sim->counts, 0, sizeof(sim->counts));
  for (size_t k = sim->pending.size(); k-- > 0;) {
    PendingMol& pm = sim->pending[k];
    int l = lookup[pm.ident][MSsoln];
    ident[l].push_back(pm.ident);
    pos[l].insert(pos[l].end(), pm.pos, pm.pos + 3);
    ser[l].push_back(pm.serno);
    sim->counts[pm.ident]++;
  }
  long ntot = (long)sim->pending.size();
  EngineConfig cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.f32 = sim->precision == 32;
  // FP32 offsets round a stored position by up to 2^-24 of the box size,
  // far more than fixpt2panel's nudges off a panel (multiples of DBL_EPSILON):
  // a molecule stored next to a reflective panel could land on its far side
  // absorbing walls without surfaces run on the surface variants with no
  // panels (checkwalls kills need the molsorts those variants carry); with
  // surfaces the walls are not applied at all (smolsim.cpp:2618-2630)
  bool wall_absorb = false;
  for (int w = 0; w < 6; w++) wall_absorb |= sim->wtype[w] == 'a';
  wall_absorb = wall_absorb && !sim->surf.enabled();
  LCHECK(!(cfg.f32 && (sim->surf.enabled() || wall_absorb)), fn, ECerror,
         "FP32 box-offset records cannot be combined with surfaces or absorbing walls (use precision 64)");
  {  // tuning override of the pair search's z sub-bins per box (1, 2, 4 or 8)
    const char* e = getenv("SMOLAMD_ZSUB_MAX");
    cfg.zsub_max = e ? atoi(e) : 0;
  }
  // bounce by reflection (rparam -2, doreact smolreact.c:2756-2762) reads the
  // reactants' positions before the step's diffusion: the engine keeps them
  for (auto& rx : sim->rxn[2])
    if (rx.rparamt == RPbounce && rx.rparam == -2) cfg.reflect = 1;
  LCHECK(!(cfg.reflect && (sim->surf.enabled() || wall_absorb || for_batch)), fn, ECerror,
         "the reflection method of bounce placement is implemented on the device path for solution models "
         "without surfaces, absorbing walls or ensembles");
  cfg.nlist = nlist;
  // a list can only hold molecules placed at setup or created by reactions;
  // lists that receive neither keep a token capacity
  std::vector<int> reachable(nlist, 0);
  for (int l = 0; l < nlist; l++) reachable[l] = !ident[l].empty();
  for (int order = 0; order <= 2; order++)
    for (auto& rx : sim->rxn[order])
      for (int p = 0; p < rx.nprod; p++) reachable[lookup[rx.prd[p]][MSsoln]] = 1;
  // with surfaces, molecules can adsorb into the lists of the bound states
  if (sim->surf.enabled())
    for (int i = 1; i < nsp; i++)
      for (int ms = MSfront; ms <= MSback; ms++)
        if (lookup[i][ms] >= 0) reachable[lookup[i][ms]] = 1;
  long capl = sim->maxmol > 0 ? sim->maxmol : ntot + (ntot / 4 > 4096 ? ntot / 4 : 4096);
  if (sim->growcap > capl) capl = sim->growcap;
  sim->devcap = capl;
  long captot = 0;
  for (int l = 0; l < nlist; l++) {
    cfg.cap[l] = reachable[l] ? (int)capl : 256;
    captot += cfg.cap[l];
  }
  LCHECK(captot < (1L << 27), fn, ECbounds, "too many molecules for the device key layout");
  {
    long maxl = 1;
    for (int l = 0; l < nlist; l++) maxl = (long)ident[l].size() > maxl ? (long)ident[l].size() : maxl;
    cfg.expected_per_box = (double)maxl / (double)g.nbox;
  }
  if (!for_batch) {
    cfg.maxedges = (int)(captot / 2 > 65536 ? captot / 2 : 65536);
    cfg.maxevents = (int)(captot / 16 > 16384 ? captot / 16 : 16384);
  } else {  // one of many: event buffers in proportion to the member's size
    cfg.maxedges = (int)(captot > 4096 ? captot : 4096);
    cfg.maxevents = (int)(captot / 4 > 2048 ? captot / 4 : 2048);
  }
  // surfaces (surfupdate, smolsurface.c:3797-3893, and the box panel lists)
  SurfaceTables st;
  HostSurf hs;
  memset(&hs, 0, sizeof(hs));
  if (sim->surf.enabled()) {
    bool mob[kMaxList] = {false};
    for (int i = 1; i < nsp; i++)
      for (int ms = 0; ms < kMS; ms++)
        if (ismobile(sim, i, ms) && lookup[i][ms] >= 0) mob[lookup[i][ms]] = true;
    std::vector<std::array<double, kMS>> dms(sim->difc.begin(), sim->difc.begin() + nsp);
    std::vector<std::array<int, kMS>> lms(lookup.begin(), lookup.begin() + nsp);
    char msg[400] = "";
    int er = sim->surf.build(g, nsp, reinterpret_cast<const double(*)[kMS]>(dms.data()), sim->dt,
                             reinterpret_cast<const int(*)[kMS]>(lms.data()), mob, nlist, st, msg, sizeof(msg));
    LCHECK(er == 0, fn, ECerror, "%s", msg);
    for (int ll = 0; ll < nlist; ll++)
      LCHECK(!(st.listbound[ll] && mob[ll]), fn, ECerror,
             "a molecule list holds both diffusing and surface-bound states; the device path needs them apart");
    // reactions of surface-bound molecules are not on the device path
    for (int order = 1; order <= 2; order++)
      for (auto& rx : sim->rxn[order])
        for (int k = 0; k < (order == 1 ? kMS : kMS1 * kMS1); k++) {
          const int ms1 = order == 1 ? k : k / kMS1, ms2 = order == 1 ? 0 : k % kMS1;
          const bool soln1 = ms1 == MSsoln || ms1 == MSbsoln, soln2 = ms2 == MSsoln || ms2 == MSbsoln;
          LCHECK(!rx.permit[k] || (soln1 && soln2), fn, ECerror,
                 "reactions of surface-bound molecules are not implemented on the device path");
        }
    hs.actb = st.actb.data();
    for (int ll = 0; ll < kMaxList; ll++) hs.listbound[ll] = st.listbound[ll];
    for (int i = 0; i < kMaxSpecies; i++) {
      hs.blist[i][0] = st.blist[i][0];
      hs.blist[i][1] = st.blist[i][1];
    }
    hs.on = 1;
    hs.npnl = (int)st.pnl.size();
    hs.pnl = st.pnl.data();
    hs.nbox = g.nbox;
    hs.bpstart = st.bpstart.data();
    hs.bplist = st.bplist.data();
    hs.nsrf = (int)sim->surf.srf.size();
    hs.act = st.act.data();
    for (int ll = 0; ll < kMaxList; ll++) hs.listdiffuse[ll] = st.listdiffuse[ll];
    hs.epsilon = sim->surf.epsilon;
  }
  std::vector<int> wall_bpstart;  // absorbing walls: every box lists no panel
  if (wall_absorb) {
    wall_bpstart.assign((size_t)g.nbox + 1, 0);
    hs.on = 1;
    hs.walls = 1;
    hs.nbox = g.nbox;
    hs.bpstart = wall_bpstart.data();
    for (int i = 0; i < kMaxSpecies; i++) hs.blist[i][0] = hs.blist[i][1] = -1;
  }
  if (for_batch) {
    LCHECK(!sim->surf.enabled() && !wall_absorb, fn, ECerror,
           "ensemble members cannot have surfaces or absorbing walls");
    sim->prep.cfg = cfg;
    sim->prep.tab = t;
    sim->prep.geo = g;
    sim->prep.ident = ident;
    sim->prep.pos = pos;
    sim->prep.ser = ser;
    return ECok;
  }
  if (sim->dev) engine_destroy(sim->dev);
  char err[256] = "";
  sim->dev = engine_create(cfg, t, g, sim->gtable.data(), 4096, (uint32_t)sim->key, (uint32_t)(sim->key >> 32), &hs,
                           err, sizeof(err));
  LCHECK(sim->dev, fn, ECmemory, "device engine: %s", err);
  for (int l = 0; l < nlist; l++) {
    int er = engine_upload_list(sim->dev, l, (int)ident[l].size(), ident[l].data(), pos[l].data(), ser[l].data());
    LCHECK(er == 0, fn, ECmemory, "upload of list %d failed", l);
  }
  engine_set_counters(sim->dev, sim->serno, sim->stepindex, sim->counts);
  sim->ok = true;
  return ECok;
}

ErrorCode engine_error(const char* fn, int bits) {
  if (bits == 0) return ECok;
  if (bits == -2) {
    set_error(fn, ECbounds, "fewer injected displacements than diffusing molecules");
    return ECbounds;
  }
  if (bits < 0) {
    set_error(fn, ECerror, "HIP runtime error");
    return ECerror;
  }
  static const char* const names[] = {"candidate pair space exceeded", "event space exceeded",
                                       "molecule list capacity exceeded", "sort space exceeded",
                                       "surface outcome not run on the device", "stepping halted"};
  std::string msg;
  for (int k = 0; k < 6; k++)
    if (bits & (1 << k)) {
      if (!msg.empty()) msg += ", ";
      msg += names[k];
    }
  set_error(fn, ECerror, "device engine error: %s (bits 0x%x)", msg.c_str(), bits);
  return ECerror;
}

// one simulatetimestep; returns the reference's code (0, 1 at tmax, 10 at tbreak) or <0
bool rxnlog_on(simptr sim) {
  for (int o = 0; o < 3; o++)
    for (auto& r : sim->rxn[o])
      if (!r.logfile.empty()) return true;
  return false;
}

// The reaction log lines of the step that started at `time` (doreact,
// smolreact.c:2879-2896): for each event of a logged reaction, in creation
// order, the line starts at the first product whose serial number (its low
// 32 bits, or its high ones for a compound serial) is listed, printing "?"
// for the products before it: time, reaction name, reaction position,
// reactant serials, then the product serials.
void write_rxnlog(simptr sim, double time) {
  std::vector<LogRow> rows;
  int n = engine_get_log(sim->dev, nullptr, 0);
  if (n <= 0) return;
  rows.resize((size_t)n);
  engine_get_log(sim->dev, rows.data(), n);
  auto listed = [](const Rxn& rx, uint64_t s) {
    if (rx.logall) return true;
    for (long v : rx.logserno)
      if (v == (long)(s & 0xFFFFFFFFull) || (s > 0xFFFFFFFull && v == (long)(s >> 32))) return true;
    return false;
  };
  for (const LogRow& row : rows) {
    if (row.order < 0 || row.order > 2 || row.r < 0 || row.r >= (int)sim->rxn[row.order].size()) continue;
    const Rxn& rx = sim->rxn[row.order][(size_t)row.r];
    if (rx.logfile.empty()) continue;
    int first = -1;
    for (int q = 0; q < rx.nprod && q < kMaxProd; q++)
      if (listed(rx, row.prd[q])) {
        first = q;
        break;
      }
    if (first < 0) continue;
    FILE* fp = nullptr;
    int data = -1;
    if (cmd::target(sim->cmds, rx.logfile.c_str(), 1, &fp, &data) < 0 || !fp) continue;
    cmd::write_rxnlog_line(sim->cmds, fp, time, rx.name.c_str(), row.pos, sim->dim, row.order, row.rct, first,
                           rx.nprod, row.prd);
  }
}

// Lists that fill up grow (molexpandlist, smolmolec.c:1911-1950, doubles a
// full list).  The device's lists have a fixed capacity, so when a list is
// more than half full after a step the engine is rebuilt from its own state
// with twice the capacity: the same virtual boxes, the molecules in their
// live order with their serial numbers, the serial counter and the step
// index.  Every draw is keyed by serial number and step, so the steps that
// follow are the ones the old engine would have taken.  Not for models with
// surfaces or absorbing walls (their molecules' surface states are not
// carried) nor under a maximum set by smolSetMaxMolecules.
ErrorCode grow_lists(simptr sim, const char* fn) {
  if (!sim->dev || sim->maxmol > 0 || sim->surf.enabled()) return ECok;
  for (int w = 0; w < 6; w++)
    if (sim->wtype[w] == 'a') return ECok;
  Control c;
  LCHECK(engine_sync_control(sim->dev, &c) == 0, fn, ECerror, "HIP runtime error");
  const int nlist = (int)sim->listname.size();
  bool full = false;
  for (int l = 0; l < nlist; l++) full |= 2L * c.nl[l] > sim->devcap;
  if (!full) return ECok;
  // the live lists, each pushed last molecule first: the update reads the
  // pending molecules from the end (the order molsort gives newest first)
  std::vector<PendingMol> pend;
  for (int l = 0; l < nlist; l++) {
    const int n = c.nl[l];
    std::vector<int> id((size_t)n), box((size_t)n);
    std::vector<double> pos((size_t)n * 3);
    std::vector<uint64_t> ser((size_t)n);
    LCHECK(engine_get_list(sim->dev, l, id.data(), pos.data(), ser.data(), box.data(), n) == n, fn, ECerror,
           "HIP runtime error");
    for (int k = n - 1; k >= 0; k--) {
      PendingMol pm;
      pm.ident = id[(size_t)k];
      pm.pos[0] = pos[3 * (size_t)k];
      pm.pos[1] = pos[3 * (size_t)k + 1];
      pm.pos[2] = pos[3 * (size_t)k + 2];
      pm.serno = ser[(size_t)k];
      pend.push_back(pm);
    }
  }
  sim->molsteps_base += c.molsteps;
  const int nlog = engine_get_log(sim->dev, nullptr, 0);
  sim->heldlog.resize(nlog > 0 ? (size_t)nlog : 0);
  if (nlog > 0) engine_get_log(sim->dev, sim->heldlog.data(), nlog);
  sim->holdlog = true;
  sim->pending.swap(pend);
  sim->serno = c.serno_next;
  sim->growcap = 2 * sim->devcap;
  sim->ok = false;
  const ErrorCode ec = do_update(sim, fn, false, true);
  if (ec == ECok) sim->rebuilds++;
  return ec;
}

int run_one(simptr sim, const double* inject, long ninject, const char* fn, ErrorCode* ec) {
  *ec = do_update(sim, fn);
  if (*ec != ECok) return -1;
  sim->started = true;
  sim->holdlog = false;
  int bits = engine_step(sim->dev, inject, ninject);
  if (bits) {
    *ec = engine_error(fn, bits);
    return -1;
  }
  if (rxnlog_on(sim)) write_rxnlog(sim, sim->time);
  sim->time += sim->dt;
  sim->stepindex++;
  *ec = grow_lists(sim, fn);
  if (*ec != ECok) return -1;
  if (sim->time >= sim->tmax) return 1;
  if (sim->time >= sim->tbreak) return 10;
  return 0;
}

// Write the queued command outputs: wait for the device, read the
// observation rows, format them.  Returns the device error bits.
int flush_outputs(simptr sim) {
  int bits = sim->dev ? engine_finish(sim->dev) : 0;
  std::vector<cmd::ObsRow> rows;
  if (sim->obsrows > 0 && sim->dev) {
    std::vector<int> buf((size_t)sim->obsrows * kObsStride);
    if (engine_read_obs(sim->dev, sim->obsrows, buf.data()) == 0) {
      rows.resize((size_t)sim->obsrows);
      for (int r = 0; r < sim->obsrows; r++) {
        const int* o = buf.data() + (size_t)r * kObsStride;
        rows[(size_t)r].step = (uint32_t)o[0];
        rows[(size_t)r].count.assign(o + 1, o + 1 + kMaxSpecies);
        rows[(size_t)r].countf.assign(o + 1 + kMaxSpecies, o + 1 + 2 * kMaxSpecies);
        rows[(size_t)r].countk.assign(o + 1 + 2 * kMaxSpecies, o + 1 + 3 * kMaxSpecies);
      }
    } else if (!bits) {
      bits = -1;
    }
  }
  cmd::flush(sim->cmds, rows, sim->spname);
  sim->obsrows = 0;
  return bits;
}

// Every molecule for listmols, in molscancmd's order (smolmolec.c:1189-1205):
// lists in order, live order within a list.  Queued outputs are written first.
int dump_molecules(simptr sim, std::vector<cmd::MolRow>* out) {
  out->clear();
  if (sim->dev && sim->ok) {
    flush_outputs(sim);
    const int nl = (int)sim->listname.size();
    for (int ll = 0; ll < nl; ll++) {
      int n = engine_get_list_inst(sim->dev, 0, ll, nullptr, nullptr, nullptr, nullptr, 0);
      if (n < 0) return 1;
      if (n == 0) continue;
      std::vector<int> id((size_t)n), ms((size_t)n);
      std::vector<double> pos((size_t)n * 3);
      std::vector<uint64_t> ser((size_t)n);
      if (engine_get_list_inst(sim->dev, 0, ll, id.data(), pos.data(), ser.data(), nullptr, n, ms.data()) < 0)
        return 1;
      for (int k = 0; k < n; k++) {
        cmd::MolRow m;
        m.ident = id[(size_t)k];
        m.ms = ms[(size_t)k];
        for (int d = 0; d < 3; d++) m.pos[d] = pos[(size_t)k * 3 + (size_t)d];
        m.serno = ser[(size_t)k];
        out->push_back(m);
      }
    }
    return 0;
  }
  // before the first update: the molecules added so far, list by list
  for (int ll = -1; ll < (int)sim->listname.size(); ll++)
    for (auto& pm : sim->pending) {
      const int l2 = sim->listlookup[(size_t)pm.ident][MSsoln];
      if (l2 != ll) continue;
      cmd::MolRow m;
      m.ident = pm.ident;
      for (int d = 0; d < 3; d++) m.pos[d] = pm.pos[d];
      m.serno = pm.serno;
      out->push_back(m);
    }
  return 0;
}

// smolsimulate on the device (smolsim.cpp:2747-2760): commands run between
// steps; an observation enqueues a copy of the species counters on the
// stream, so the steps are never waited for except every 256 steps (error
// check) and when the observation table is full.  Returns the reference's
// code (1 tmax, 7 stopped by a command, 10 tbreak) or -1 after a device error
// (*ec set).
int simulate_device(simptr sim, const char* fn, ErrorCode* ec) {
  *ec = do_update(sim, fn);
  if (*ec != ECok) return -1;
  sim->started = true;
  cmd::update(sim->cmds, sim->tmin, sim->tmax, sim->dt);
  cmd::Clock k;
  k.tmin = sim->tmin;
  k.tmax = sim->tmax;
  k.dt = sim->dt;
  k.time = sim->time;
  k.tbreak = sim->tbreak;
  const double t0 = sim->time;
  const uint32_t s0 = sim->stepindex;
  int bits = 0;
  long since = 0, sincegrow = 0;
  ErrorCode growerr = ECok;
  auto observe = [&](uint32_t* step) -> int {
    if (sim->obsrows >= kObsRows) {
      int b = flush_outputs(sim);
      if (b && !bits) bits = b;
    }
    if (engine_observe(sim->dev, sim->obsrows)) return -1;
    *step = sim->stepindex;
    return sim->obsrows++;
  };
  auto dump = [&](std::vector<cmd::MolRow>* mols) -> int {
    int b = flush_outputs(sim);
    if (b && !bits) bits = b;
    return bits ? 1 : dump_molecules(sim, mols);
  };
  auto ex = [&](cmd::Command& c, double time) -> int {
    std::string err;
    int code = cmd::run_command(sim->cmds, c, time, sim->spname, observe, dump, &err);
    if (code == cmd::kWarn) fprintf(stderr, "command '%s' error: %s\n", c.str.c_str(), err.c_str());
    return code;
  };
  const bool logging = rxnlog_on(sim);
  double tstep = sim->time;  // start time of the next step, accumulated as the clock does
  auto step = [&]() -> int {
    if (bits) return -1;
    sim->holdlog = false;
    if (engine_enqueue(sim->dev, 1)) {
      bits = -1;
      return -1;
    }
    sim->stepindex++;
    if (logging) {  // the step's events are written before the commands that follow it
      int b = flush_outputs(sim);
      if (b) {
        bits = b;
        return -1;
      }
      write_rxnlog(sim, tstep);
    }
    tstep += sim->dt;
    if (++since >= 256) {
      since = 0;
      int b = flush_outputs(sim);
      if (b) {
        bits = b;
        return -1;
      }
    }
    if (++sincegrow >= kGrowCheckSteps) {  // full lists grow between steps
      sincegrow = 0;
      int b = flush_outputs(sim);
      if (b) {
        bits = b;
        return -1;
      }
      growerr = grow_lists(sim, fn);
      if (growerr != ECok) return -1;
    }
    return 0;
  };
  int code = cmd::simulate(sim->cmds, k, step, ex);
  int b = flush_outputs(sim);
  if (b && !bits) bits = b;
  sim->time = k.time;
  if (growerr == ECok && !bits) growerr = grow_lists(sim, fn);  // the device is idle after the flush
  if (growerr != ECok) {
    *ec = growerr;
    return -1;
  }
  if (bits) {
    // the device stops at the failing step: the clock follows its step counter
    Control c;
    if (engine_sync_control(sim->dev, &c) == 0) {
      sim->stepindex = c.step;
      double t = t0;
      for (uint32_t q = s0; q != c.step; q++) t += sim->dt;
      sim->time = t;
    }
    *ec = engine_error(fn, bits);
    return -1;
  }
  return code;
}

}  // namespace

// ===========================================================================
// exported API

extern "C" {

double smolGetVersion(void) { return 2.76; }

void smolSetError(const char* errorfunction, enum ErrorCode errorcode, const char* errorstring, const char* flags) {
  (void)flags;
  if (errorcode != ECsame) {
    g_errcode = errorcode;
    g_warncode = errorcode >= ECwarning ? errorcode : ECok;
    if (errorstring) {
      strncpy(g_errstr, errorstring, kStrLen - 1);
      g_errstr[kStrLen - 1] = 0;
    } else
      g_errstr[0] = 0;
  }
  if (errorfunction) {
    strncpy(g_errfunc, errorfunction, kStrLen - 1);
    g_errfunc[kStrLen - 1] = 0;
  } else
    g_errfunc[0] = 0;
  if (g_debugmode && g_errfunc[0] && g_errcode < ECwarning) {
    char s[64];
    fprintf(stderr, "Libsmoldyn '%s' error in %s: %s\n", smolErrorCodeToString(g_errcode, s), g_errfunc, g_errstr);
  }
}

enum ErrorCode smolGetError(char* errorfunction, char* errorstring, int clearerror) {
  ErrorCode erc = g_errcode;
  if (errorfunction) strcpy(errorfunction, g_errfunc);
  if (errorstring) strcpy(errorstring, g_errstr);
  if (clearerror) smolClearError();
  return erc;
}

void smolClearError(void) {
  g_errcode = ECok;
  g_warncode = ECok;
  g_errfunc[0] = 0;
  g_errstr[0] = 0;
}

void smolSetDebugMode(int debugmode) { g_debugmode = debugmode; }

char* smolErrorCodeToString(enum ErrorCode erc, char* string) {
  static const char* names[] = {"ok", "notify", "warning", "nonexistent", "all", "missing", "bounds",
                                "syntax", "error", "memory", "bug", "same", "wildcard"};
  int k = -(int)erc;
  strcpy(string, (k >= 0 && k <= 12) ? names[k] : "undefined");
  return string;
}

simptr smolNewSim(int dim, double* lowbounds, double* highbounds) {
  const char* fn = "smolNewSim";
  if (dim <= 0 || dim > 3) {
    set_error(fn, ECbounds, "dim must be between 1 and 3");
    return nullptr;
  }
  if (!lowbounds || !highbounds) {
    set_error(fn, ECmissing, "missing bounds");
    return nullptr;
  }
  for (int d = 0; d < dim; d++)
    if (!(lowbounds[d] < highbounds[d])) {
      set_error(fn, ECbounds, "lowbounds must be < highbounds");
      return nullptr;
    }
  simptr sim = new smolamd_sim();
  sim->dim = dim;
  for (int d = 0; d < dim; d++) {
    sim->low[d] = lowbounds[d];
    sim->high[d] = highbounds[d];
  }
  return sim;
}

enum ErrorCode smolUpdateSim(simptr sim) {
  const char* fn = "smolUpdateSim";
  LCHECK(sim, fn, ECmissing, "missing sim");
  return do_update(sim, fn);
}

// libsmoldyn.cpp:217-238: one step through smolsimulate with the break time
// half a step ahead (the break time stays set, as in the reference)
enum ErrorCode smolRunTimeStep(simptr sim) {
  const char* fn = "smolRunTimeStep";
  LCHECK(sim, fn, ECmissing, "missing sim");
  sim->tbreak = sim->time + sim->dt / 2;
  ErrorCode ec;
  int code = simulate_device(sim, fn, &ec);
  if (code < 0) return ec;
  LCHECK(code != 1, fn, ECnotify, "Simulation complete");
  LCHECK(code != 7, fn, ECnotify, "Simulation stopped by a runtime command");
  return g_warncode;
}

// libsmoldyn.cpp:242-270: output files are opened (overwriting) first
enum ErrorCode smolRunSim(simptr sim) {
  const char* fn = "smolRunSim";
  LCHECK(sim, fn, ECmissing, "missing sim");
  char msg[kStrLen];
  LCHECK(cmd::open_files(sim->cmds, true, msg, (int)sizeof(msg)) == 0, fn, ECerror,
         "Cannot open output files for writing");
  ErrorCode ec;
  int code = simulate_device(sim, fn, &ec);
  if (code < 0) return ec;
  LCHECK(code != 1, fn, ECnotify, "Simulation complete");
  LCHECK(code != 7, fn, ECnotify, "Simulation stopped by a runtime command");
  if (g_warncode == ECnotify) g_warncode = ECok;
  return g_warncode;
}

// libsmoldyn.cpp:274-284: the break time is the stop time afterwards
enum ErrorCode smolRunSimUntil(simptr sim, double breaktime) {
  const char* fn = "smolRunSimUntil";
  LCHECK(sim, fn, ECmissing, "missing sim");
  const double stoptime = sim->tmax;
  sim->tbreak = breaktime;
  smolRunSim(sim);
  sim->tbreak = stoptime;
  return g_warncode;
}

enum ErrorCode smolFreeSim(simptr sim) {
  if (sim) {
    if (sim->dev) engine_destroy(sim->dev);
    delete sim;
  }
  return ECok;
}

enum ErrorCode smolSetSimTimes(simptr sim, double timestart, double timestop, double timestep) {
  const char* fn = "smolSetSimTimes";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(timestep > 0, fn, ECbounds, "timestep must be >0");
  sim->tmin = sim->time = timestart;
  sim->tmax = timestop;
  sim->dt = timestep;
  cmd::set_condition(sim->cmds, 1, 0);
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolSetTimeStart(simptr sim, double timestart) {
  const char* fn = "smolSetTimeStart";
  LCHECK(sim, fn, ECmissing, "missing sim");
  sim->tmin = timestart;
  return ECok;
}

enum ErrorCode smolSetTimeStop(simptr sim, double timestop) {
  const char* fn = "smolSetTimeStop";
  LCHECK(sim, fn, ECmissing, "missing sim");
  sim->tmax = timestop;
  return ECok;
}

enum ErrorCode smolSetTimeNow(simptr sim, double timenow) {
  const char* fn = "smolSetTimeNow";
  LCHECK(sim, fn, ECmissing, "missing sim");
  sim->time = timenow;
  return ECok;
}

enum ErrorCode smolSetTimeStep(simptr sim, double timestep) {
  const char* fn = "smolSetTimeStep";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(timestep > 0, fn, ECbounds, "timestep must be >0");
  LCHECK(!sim->started, fn, ECerror, "the time step is fixed once stepping started");
  sim->dt = timestep;
  cmd::set_condition(sim->cmds, 1, 0);
  invalidate(sim);
  return ECok;
}

// Record precision of the device path (no counterpart in the reference):
// 64 keeps FP64 positions; 32 stores each position as FP32 offsets from the
// FP64 origin of its virtual box, with all arithmetic (diffusion, walls,
// surfaces, distance tests, product placement) still in FP64.
enum ErrorCode smolamdSetPrecision(simptr sim, int bits) {
  const char* fn = "smolamdSetPrecision";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(bits == 32 || bits == 64, fn, ECbounds, "precision must be 32 or 64");
  LCHECK(!sim->started, fn, ECerror, "the precision is fixed once stepping started");
  sim->precision = bits;
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolSetRandomSeed(simptr sim, long int seed) {
  const char* fn = "smolSetRandomSeed";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(!sim->started, fn, ECerror, "the seed is fixed once stepping started");
  sim->key = (uint64_t)seed;
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolSetPartitions(simptr sim, const char* method, double value) {
  const char* fn = "smolSetPartitions";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(method, fn, ECmissing, "missing method");
  LCHECK(value > 0, fn, ECbounds, "value must be >0");
  if (!strcmp(method, "molperbox")) sim->mpbox = value;
  else if (!strcmp(method, "boxsize")) sim->boxsize = value;
  else LCHECK(0, fn, ECsyntax, "method not recognized");
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolAddSpecies(simptr sim, const char* species, const char* mollist) {
  const char* fn = "smolAddSpecies";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(species && species[0], fn, ECmissing, "missing species");
  LCHECK(species_index(sim, species) < 0, fn, ECerror, "species is already in system");
  LCHECK(strcmp(species, "all") != 0 && strcmp(species, "empty") != 0, fn, ECsyntax, "species name not allowed");
  LCHECK((int)sim->spname.size() < kMaxSpecies, fn, ECbounds, "too many species");
  int ll = -1;
  if (mollist && mollist[0]) {
    ll = list_index(sim, mollist);
    LCHECK(ll >= 0, fn, ECnonexist, "molecule list not found");
  }
  sim->spname.push_back(species);
  sim->difc.push_back({0, 0, 0, 0, 0});
  sim->listlookup.push_back({ll, ll, ll, ll, ll});
  invalidate(sim);
  return ECok;
}

// smolGetSpeciesIndexNT (libsmoldyn.cpp:851-866): the index, or the error
// code, recorded with the reference's messages and no debug report
int smolGetSpeciesIndexNT(simptr sim, const char* species) {
  const char* fn = "smolGetSpeciesIndexNT";
  if (!sim) return set_error_nt(fn, ECmissing, "missing sim");
  if (!species) return set_error_nt(fn, ECmissing, "missing species name");
  if (sim->spname.size() <= 1) return set_error_nt(fn, ECnonexist, "no species defined");
  if (!strcmp(species, "all")) return set_error_nt(fn, ECall, "species cannot be 'all'");
  int i = species_index(sim, species);
  if (i <= 0) {
    char buf[kStrLen];
    snprintf(buf, sizeof(buf), "species '%s' not found", species);
    return set_error_nt(fn, ECnonexist, buf);
  }
  return i;
}

// smolGetSpeciesIndex (libsmoldyn.cpp:834-848): the NT lookup's error,
// named for this function
int smolGetSpeciesIndex(simptr sim, const char* species) {
  int i = smolGetSpeciesIndexNT(sim, species);
  if (i <= 0) return (int)same_error("smolGetSpeciesIndex");
  return i;
}

void smolGetSpeciesName(simptr sim, int speciesindex, char* species) {
  if (!sim || !species) return;
  if (speciesindex < 0 || speciesindex >= (int)sim->spname.size()) {
    species[0] = 0;
    smolSetError("smolGetSpeciesName", ECbounds, "invalid species index", "");
    return;
  }
  strcpy(species, sim->spname[speciesindex].c_str());
}

enum ErrorCode smolSetSpeciesMobility(simptr sim, const char* species, enum MolecState state, double difc,
                                      double* drift, double* difmatrix) {
  const char* fn = "smolSetSpeciesMobility";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(difc >= 0, fn, ECbounds, "negative diffusion coefficient");
  LCHECK(!drift && !difmatrix, fn, ECerror, "drift and anisotropic diffusion are not implemented on the device path");
  LCHECK(!sim->started, fn, ECerror, "mobility is fixed once stepping started");
  std::vector<int> idx;
  const int isp = smolGetSpeciesIndexNT(sim, species);
  if (isp == (int)ECall) {
    smolClearError();
    for (size_t i = 1; i < sim->spname.size(); i++) idx.push_back((int)i);
  } else {
    if (isp <= 0) return same_error(fn);
    idx.push_back(isp);
  }
  for (int i : idx) {
    if (state == MSall)
      for (int ms = 0; ms < kMS; ms++) sim->difc[i][ms] = difc;
    else if (state == MSbsoln || state == MSsoln)
      sim->difc[i][MSsoln] = difc;
    else if (state < kMS)
      sim->difc[i][state] = difc;
    else
      LCHECK(0, fn, ECbounds, "invalid state");
  }
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolAddMolList(simptr sim, const char* mollist) {
  const char* fn = "smolAddMolList";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(mollist && mollist[0], fn, ECmissing, "missing mollist");
  LCHECK(list_index(sim, mollist) < 0, fn, ECerror, "molecule list already exists");
  LCHECK((int)sim->listname.size() < kMaxList, fn, ECbounds, "too many molecule lists");
  sim->listname.push_back(mollist);
  invalidate(sim);
  return ECok;
}

// smolGetMolListIndexNT (libsmoldyn.cpp:987-999)
int smolGetMolListIndexNT(simptr sim, const char* mollist) {
  const char* fn = "smolGetMolListIndexNT";
  if (!sim) return set_error_nt(fn, ECmissing, "missing sim");
  if (!mollist) return set_error_nt(fn, ECmissing, "missing mollist");
  if (sim->listname.empty()) return set_error_nt(fn, ECnonexist, "no molecule lists defined");
  if (!strcmp(mollist, "all")) return set_error_nt(fn, ECall, "molecule list cannot be 'all'");
  int l = list_index(sim, mollist);
  if (l < 0) return set_error_nt(fn, ECnonexist, "list name not recognized");
  return l;
}

int smolGetMolListIndex(simptr sim, const char* mollist) {
  int l = smolGetMolListIndexNT(sim, mollist);
  if (l < 0) return (int)same_error("smolGetMolListIndex");
  return l;
}

enum ErrorCode smolSetMolList(simptr sim, const char* species, enum MolecState state, const char* mollist) {
  const char* fn = "smolSetMolList";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int i = smolGetSpeciesIndexNT(sim, species);
  if (i == (int)ECall) {
    smolClearError();
    i = -5;  // every species
  } else if (i <= 0) {
    return same_error(fn);
  }
  LCHECK((state >= 0 && state < kMS) || state == MSbsoln || state == MSall, fn, ECsyntax, "invalid state");
  int ll = smolGetMolListIndexNT(sim, mollist);
  if (ll < 0) return same_error(fn);
  for (int k = (i == -5 ? 1 : i); k < (i == -5 ? (int)sim->spname.size() : i + 1); k++) {
    if (state == MSall)
      for (int ms = 0; ms < kMS; ms++) sim->listlookup[k][ms] = ll;
    else
      sim->listlookup[k][state == MSbsoln ? MSsoln : state] = ll;
  }
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolSetMaxMolecules(simptr sim, int maxmolecules) {
  const char* fn = "smolSetMaxMolecules";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(maxmolecules > 0, fn, ECbounds, "maxmolecules must be >0");
  sim->maxmol = maxmolecules;
  invalidate(sim);
  return ECok;
}

// Molecules added once stepping started (addmol with sort 0,
// smolmolec.c:2717-2740): the reference resurrects them and runs no molsort,
// so they join their list at the first molsort of the next step, after
// diffusion, the wall or surface checks and reassignmolecs
// (smolsim.cpp:2607-2636): appended in reverse creation order (the dead
// list's order), undiffused in that step.  Serial numbers continue the
// device's counter.  Models on the surface variants are refused (their first
// molsort would interleave the new molecules with list changes).
static ErrorCode add_running(simptr sim, const char* fn, int ident, int number, const double* lo, const double* hi,
                      const double* positions) {
  LCHECK(sim->dev, fn, ECerror, "no device engine");
  Control c;
  LCHECK(engine_get_control(sim->dev, 0, &c) == 0, fn, ECerror, "HIP runtime error");
  std::vector<PendingMol> mols((size_t)number);
  for (int k = 0; k < number; k++) {
    PendingMol& pm = mols[(size_t)k];
    pm.ident = ident;
    pm.serno = c.serno_next + (uint64_t)k;
    if (positions) {
      for (int d = 0; d < 3; d++) pm.pos[d] = positions[3 * k + d];
    } else {
      smolrng4 r = smolrng_mol((uint32_t)sim->key, (uint32_t)(sim->key >> 32), pm.serno, 0, SMOLRNG_PLACE, 0);
      for (int d = 0; d < 3; d++) pm.pos[d] = smolrng_u01open(r.v[d]) * (hi[d] - lo[d]) + lo[d];
    }
  }
  const int ll = sim->listlookup[(size_t)ident][MSsoln];
  std::vector<int> id((size_t)number, ident);
  std::vector<double> pos((size_t)number * 3);
  std::vector<uint64_t> ser((size_t)number);
  for (int k = 0; k < number; k++) {  // reverse creation order
    const PendingMol& pm = mols[(size_t)(number - 1 - k)];
    for (int d = 0; d < 3; d++) pos[(size_t)k * 3 + d] = pm.pos[d];
    ser[(size_t)k] = pm.serno;
  }
  const int er = engine_append_fresh(sim->dev, ll, number, id.data(), pos.data(), ser.data());
  LCHECK(er != -1, fn, ECerror,
         "molecules can be added after stepping started only to models without surfaces, absorbing walls, "
         "ensembles or FP32 records");
  LCHECK(er == 0, fn, ECmemory, "the molecule list is full (set max_mol)");
  return ECok;
}

enum ErrorCode smolAddSolutionMolecules(simptr sim, const char* species, int number, double* lowposition,
                                        double* highposition) {
  const char* fn = "smolAddSolutionMolecules";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int i = smolGetSpeciesIndexNT(sim, species);
  if (i <= 0) return same_error(fn);
  LCHECK(number >= 0, fn, ECbounds, "number cannot be negative");
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = lowposition ? lowposition[d] : sim->low[d];
    hi[d] = highposition ? highposition[d] : sim->high[d];
  }
  if (sim->started) return add_running(sim, fn, i, number, lo, hi, nullptr);
  // placement in the counter stream: unirandOOD per coordinate, keyed by the
  // new molecule's serial number (addmol, smolmolec.c:2717-2740)
  for (int k = 0; k < number; k++) {
    PendingMol pm;
    pm.ident = i;
    pm.serno = sim->serno++;
    smolrng4 r = smolrng_mol((uint32_t)sim->key, (uint32_t)(sim->key >> 32), pm.serno, 0, SMOLRNG_PLACE, 0);
    for (int d = 0; d < 3; d++) pm.pos[d] = smolrng_u01open(r.v[d]) * (hi[d] - lo[d]) + lo[d];
    sim->pending.push_back(pm);
  }
  invalidate(sim);
  return ECok;
}

int smolGetMoleculeCount(simptr sim, const char* species, enum MolecState state) {
  const char* fn = "smolGetMoleculeCount";
  if (!sim) {
    set_error(fn, ECmissing, "missing sim");
    return (int)ECmissing;
  }
  if (state != MSsoln && state != MSall && state != MSfront && state != MSback) return 0;
  int counts[kMaxSpecies] = {0};
  if (sim->dev && sim->ok) {
    Control c;
    if (engine_sync_control(sim->dev, &c)) {
      set_error(fn, ECerror, "HIP runtime error");
      return (int)ECerror;
    }
    // molcount by state: all states, solution = all minus the bound ones
    for (int i = 0; i < kMaxSpecies; i++)
      counts[i] = state == MSall    ? c.count[i]
                  : state == MSfront ? c.countf[i]
                  : state == MSback  ? c.countk[i]
                                     : c.count[i] - c.countf[i] - c.countk[i];
  } else if (state == MSfront || state == MSback) {
    return 0;
  } else {
    for (auto& pm : sim->pending) counts[pm.ident]++;
  }
  const int i = smolGetSpeciesIndexNT(sim, species);
  if (i == (int)ECall) {  // every species
    smolClearError();
    int s = 0;
    for (int k = 1; k < kMaxSpecies; k++) s += counts[k];
    return s;
  }
  if (i <= 0) return (int)same_error(fn);
  return counts[i];
}

// ---- configuration files (libsmoldyn.cpp:306-375)

simptr smolPrepareSimFromFile(const char* filepath, const char* filename, const char* flags) {
  const char* fn = "smolPrepareSimFromFile";
  (void)flags;
  if (!filename) {
    set_error(fn, ECmissing, "missing filename");
    return nullptr;
  }
  simptr sim = nullptr;
  std::string err;
  if (config_load(filepath, filename, &sim, &err)) {
    set_error(fn, ECerror, "%s", err.c_str());
    return nullptr;
  }
  if (do_update(sim, fn) != ECok) {
    smolFreeSim(sim);
    set_error(fn, ECerror, "Failed to update simulation");
    return nullptr;
  }
  return sim;
}

enum ErrorCode smolLoadSimFromFile(const char* filepath, const char* filename, simptr* simpointer, const char* flags) {
  const char* fn = "smolLoadSimFromFile";
  (void)flags;
  LCHECK(filename, fn, ECmissing, "missing filename");
  LCHECK(simpointer, fn, ECmissing, "missing simpointer");
  std::string err;
  LCHECK(config_load(filepath, filename, simpointer, &err) == 0, fn, ECerror, "%s", err.c_str());
  return ECok;
}

enum ErrorCode smolReadConfigString(simptr sim, const char* statement, char* parameters) {
  const char* fn = "smolReadConfigString";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(statement, fn, ECmissing, "missing statement");
  std::string cursurf, err;
  LCHECK(strcmp(statement, "start_surface") != 0, fn, ECerror,
         "Error in configuration string: surface blocks need a configuration file");
  LCHECK(config_statement(sim, statement, parameters ? parameters : "", &cursurf, &err) == 0, fn, ECerror,
         "Error in configuration string: %s", err.c_str());
  return ECok;
}

// ---- run-time commands and output (libsmoldyn.cpp:656-798)

enum ErrorCode smolSetOutputPath(simptr sim, const char* path) {
  const char* fn = "smolSetOutputPath";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(path, fn, ECmissing, "missing path");
  sim->cmds.froot = path;
  return ECok;
}

enum ErrorCode smolAddOutputFile(simptr sim, char* filename, int suffix, int append) {
  const char* fn = "smolAddOutputFile";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(filename, fn, ECmissing, "missing filename");
  LCHECK(!strchr(filename, ' '), fn, ECwarning, "only first word of filename is used");
  cmd::add_files(sim->cmds, filename, append != 0);
  if (suffix >= 0) LCHECK(!cmd::set_suffix(sim->cmds, filename, suffix), fn, ECbug, "scmdsetfsuffix bug");
  return g_warncode;
}

enum ErrorCode smolAddOutputData(simptr sim, char* dataname) {
  const char* fn = "smolAddOutputData";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(dataname, fn, ECmissing, "missing dataname");
  LCHECK(!strchr(dataname, ' '), fn, ECwarning, "only first word of dataname is used");
  cmd::add_data(sim->cmds, dataname);
  return g_warncode;
}

enum ErrorCode smolOpenOutputFiles(simptr sim, int overwrite) {
  const char* fn = "smolOpenOutputFiles";
  LCHECK(sim, fn, ECmissing, "missing sim");
  char msg[kStrLen];
  LCHECK(cmd::open_files(sim->cmds, overwrite != 0, msg, (int)sizeof(msg)) == 0, fn, ECerror,
         "Unable (or forbidden) to open output files for writing");
  return ECok;
}

enum ErrorCode smolAddCommand(simptr sim, char type, double on, double off, double step, double multiplier,
                              const char* commandstring) {
  const char* fn = "smolAddCommand";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(cmd::supported(commandstring), fn, ECerror,
         "command '%s' is not run by the device engine (stop, molcount, molcountheader, molcountspecies, "
         "molcountspecieslist, listmols, listmols2 are)",
         commandstring ? commandstring : "");
  int er = cmd::add_command(sim->cmds, type, on, off, step, multiplier, commandstring);
  LCHECK(er != 3, fn, ECsyntax, "missing command string");
  return ECok;
}

enum ErrorCode smolAddCommandFromString(simptr sim, char* string) {
  const char* fn = "smolAddCommandFromString";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(string, fn, ECmissing, "missing string");
  cmd::Commands probe;
  int er = cmd::add_command_string(probe, string);
  if (er == 0)
    LCHECK(cmd::supported(probe.list[0].c.str.c_str()), fn, ECerror,
           "command '%s' is not run by the device engine (stop, molcount, molcountheader, molcountspecies, "
           "molcountspecieslist, listmols, listmols2 are)",
           probe.list[0].c.str.c_str());
  er = cmd::add_command_string(sim->cmds, string);
  LCHECK(er != 3, fn, ECsyntax, "cmd format: type [on off dt] string");
  LCHECK(er != 5, fn, ECbounds, "cmd time step needs to be >0");
  return ECok;
}

enum ErrorCode smolGetOutputData(simptr sim, char* dataname, int* nrow, int* ncol, double** array, int erase) {
  const char* fn = "smolGetOutputData";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(dataname, fn, ECmissing, "missing dataname");
  LCHECK(nrow && ncol && array, fn, ECmissing, "missing pointer for returned data");
  LCHECK(!sim->cmds.data.empty(), fn, ECerror, "no data files in the sim");
  int did = cmd::find_data(sim->cmds, dataname);
  LCHECK(did >= 0, fn, ECerror, "no data file of the requested name");
  cmd::Commands::Data& d = sim->cmds.data[(size_t)did];
  const int nr = (int)d.rows.size(), nc = d.ncol;
  double* out = (double*)calloc((size_t)(nr * nc > 0 ? nr * nc : 1), sizeof(double));
  LCHECK(out, fn, ECmemory, "out of memory");
  for (int i = 0; i < nr; i++)
    for (size_t j = 0; j < d.rows[(size_t)i].size(); j++) out[(size_t)i * nc + j] = d.rows[(size_t)i][j];
  *nrow = nr;
  *ncol = nc;
  *array = out;
  if (erase) {
    d.rows.clear();
    d.ncol = 0;
  }
  return ECok;
}

enum ErrorCode smolRunCommand(simptr sim, const char* commandstring) {
  const char* fn = "smolRunCommand";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(commandstring, fn, ECmissing, "missing command string");
  if (sim->dev) flush_outputs(sim);
  std::vector<cmd::ObsRow> rows;
  auto observe = [&](uint32_t* step) -> int {
    cmd::ObsRow r;
    r.count.assign(kMaxSpecies, 0);
    r.countf.assign(kMaxSpecies, 0);
    r.countk.assign(kMaxSpecies, 0);
    if (sim->dev && sim->ok) {
      Control c;
      if (engine_sync_control(sim->dev, &c)) return -1;
      for (int i = 0; i < kMaxSpecies; i++) {
        r.count[(size_t)i] = c.count[i];
        r.countf[(size_t)i] = c.countf[i];
        r.countk[(size_t)i] = c.countk[i];
      }
      r.step = c.step;
    } else {
      for (auto& pm : sim->pending) r.count[(size_t)pm.ident]++;
      r.step = sim->stepindex;
    }
    *step = r.step;
    rows.push_back(r);
    return (int)rows.size() - 1;
  };
  auto dump = [&](std::vector<cmd::MolRow>* mols) -> int {
    cmd::flush(sim->cmds, rows, sim->spname);
    rows.clear();
    return dump_molecules(sim, mols);
  };
  cmd::Command c;
  c.str = commandstring;
  std::string err;
  int code = cmd::run_command(sim->cmds, c, sim->time, sim->spname, observe, dump, &err);
  cmd::flush(sim->cmds, rows, sim->spname);
  LCHECK(code == cmd::kOk, fn, ECwarning, "%s", err.empty() ? "command returned a control code" : err.c_str());
  return ECok;
}

// The command scheduler driven by a script, with no simulation attached (see
// the header for the script lines).  Each execution is an event (time,
// command-list position, iteration); each run adds an event with position
// -1000 - (its return code).  Returns the number of events, -1 on a bad line.
int smolamdCommandTimeline(const char* script, int maxev, double* ev_time, int* ev_cmd, long long* ev_iter) {
  if (!script) return -1;
  cmd::Commands cs;
  cmd::Clock k;
  int nev = 0;
  auto record = [&](double t, int c, long long it) {
    if (nev < maxev) {
      if (ev_time) ev_time[nev] = t;
      if (ev_cmd) ev_cmd[nev] = c;
      if (ev_iter) ev_iter[nev] = it;
    }
    nev++;
  };
  cmd::Exec ex = [&](cmd::Command& c, double time) -> int {
    int pos = -1;
    for (size_t i = 0; i < cs.list.size(); i++)
      if (cs.list[i].c.str == c.str) {
        pos = (int)i;
        break;
      }
    record(time, pos, cs.iter - 1);
    return c.str.compare(0, 4, "stop") == 0 ? cmd::kStop : cmd::kOk;
  };
  cmd::Step step = []() { return 0; };
  auto run = [&]() {
    cmd::update(cs, k.tmin, k.tmax, k.dt);
    int code = cmd::simulate(cs, k, step, ex);
    record(k.time, -1000 - code, cs.iter);
  };
  std::string text(script);
  size_t a = 0;
  while (a < text.size()) {
    size_t b = text.find('\n', a);
    if (b == std::string::npos) b = text.size();
    std::string line = text.substr(a, b - a);
    a = b + 1;
    char word[32] = {0};
    int nread = 0;
    if (sscanf(line.c_str(), "%31s%n", word, &nread) != 1) continue;
    const char* rest = line.c_str() + nread;
    if (!strcmp(word, "times")) {
      if (sscanf(rest, "%lf %lf %lf", &k.tmin, &k.tmax, &k.dt) != 3) return -1;
      k.time = k.tmin;
      cmd::set_condition(cs, 1, 0);
    } else if (!strcmp(word, "cmd")) {
      char t = 0;
      double on, off, st, mu;
      int n2 = 0;
      if (sscanf(rest, " %c %lf %lf %lf %lf %n", &t, &on, &off, &st, &mu, &n2) != 5) return -1;
      cmd::add_command(cs, t, on, off, st, mu, rest + n2);
    } else if (!strcmp(word, "cmdstr")) {
      cmd::add_command_string(cs, rest);
    } else if (!strcmp(word, "dt")) {
      if (sscanf(rest, "%lf", &k.dt) != 1) return -1;
      cmd::set_condition(cs, 1, 0);
    } else if (!strcmp(word, "tmax")) {
      if (sscanf(rest, "%lf", &k.tmax) != 1) return -1;
    } else if (!strcmp(word, "run")) {
      run();
    } else if (!strcmp(word, "runstep")) {
      k.tbreak = k.time + k.dt / 2;
      run();
    } else if (!strcmp(word, "rununtil")) {
      double t;
      if (sscanf(rest, "%lf", &t) != 1) return -1;
      const double stop = k.tmax;
      k.tbreak = t;
      run();
      k.tbreak = stop;
    } else {
      return -1;
    }
  }
  return nev;
}

enum ErrorCode smolSetBoundaryType(simptr sim, int dimension, int highside, char type) {
  const char* fn = "smolSetBoundaryType";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(dimension < sim->dim, fn, ECbounds, "dimension cannot exceed system dimensionality");
  LCHECK(highside < 2, fn, ECbounds, "highside must be -1, 0, or 1");
  LCHECK(type == 'r' || type == 'p' || type == 'a' || type == 't', fn, ECsyntax, "invalid type");
  for (int d = 0; d < sim->dim; d++) {
    if (dimension >= 0 && d != dimension) continue;
    for (int side = 0; side < 2; side++) {
      if (highside >= 0 && side != highside) continue;
      sim->wtype[2 * d + side] = type;
    }
  }
  invalidate(sim);
  return ECok;
}

// ---------------------------------------------------------------------------
// surfaces (libsmoldyn.cpp:1240-1480)

enum ErrorCode smolAddSurface(simptr sim, const char* surface) {
  const char* fn = "smolAddSurface";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(surface && surface[0], fn, ECmissing, "missing surface name");
  LCHECK(strcmp(surface, "all"), fn, ECall, "surface cannot be 'all'");
  LCHECK(sim->surf.find_surface(surface) < 0, fn, ECerror, "surface is already in system");
  LCHECK((int)sim->surf.srf.size() < kMaxSurf, fn, ECmemory, "too many surfaces");
  sim->surf.add_surface(surface);
  invalidate(sim);
  return ECok;
}

// smolGetSurfaceIndexNT (libsmoldyn.cpp:1274-1286)
int smolGetSurfaceIndexNT(simptr sim, const char* surface) {
  const char* fn = "smolGetSurfaceIndexNT";
  if (!sim) return set_error_nt(fn, ECmissing, "missing sim");
  if (!surface) return set_error_nt(fn, ECmissing, "missing surface");
  if (sim->surf.srf.empty()) return set_error_nt(fn, ECnonexist, "no surfaces defined");
  if (!strcmp(surface, "all")) return set_error_nt(fn, ECall, "surface cannot be 'all'");
  int s = sim->surf.find_surface(surface);
  if (s < 0) return set_error_nt(fn, ECnonexist, "surface not found");
  return s;
}

int smolGetSurfaceIndex(simptr sim, const char* surface) {
  int s = smolGetSurfaceIndexNT(sim, surface);
  if (s < 0) return (int)same_error("smolGetSurfaceIndex");
  return s;
}

// the reference passes no new species to surfsetaction here (libsmoldyn.cpp:1322-1330)
enum ErrorCode smolSetSurfaceAction(simptr sim, const char* surface, enum PanelFace face, const char* species,
                                    enum MolecState state, enum SrfAction action, const char* newspecies) {
  const char* fn = "smolSetSurfaceAction";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int s = smolGetSurfaceIndexNT(sim, surface);
  if (s == (int)ECall) smolClearError();
  else if (s < 0) return same_error(fn);
  LCHECK(face == PFfront || face == PFback || face == PFboth, fn, ECbounds, "invalid face");
  int i = smolGetSpeciesIndexNT(sim, species);
  if (i == (int)ECall) {
    smolClearError();
    i = -5;
  } else if (i <= 0) {
    return same_error(fn);
  }
  LCHECK((state >= 0 && state < kMS) || state == MSall, fn, ECbounds, "invalid state");
  LCHECK(action >= 0 && action <= SAmult, fn, ECbounds, "invalid action");
  if (newspecies && newspecies[0])
    LCHECK(smolGetSpeciesIndexNT(sim, newspecies) > 0, fn, ECnonexist, "unrecognized new species name");
  int nsp = (int)sim->spname.size();
  for (int k = 0; k < (int)sim->surf.srf.size(); k++)
    if (s < 0 || k == s) sim->surf.set_action(k, i, nsp, state, face, action);
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolSetSurfaceRate(simptr sim, const char* surface, const char* species, enum MolecState state,
                                  enum MolecState state1, enum MolecState state2, double rate, const char* newspecies,
                                  int isinternal) {
  const char* fn = "smolSetSurfaceRate";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int s = smolGetSurfaceIndexNT(sim, surface);
  if (s == (int)ECall) smolClearError();
  else if (s < 0) return same_error(fn);
  int i = smolGetSpeciesIndexNT(sim, species);
  if (i == (int)ECall) {
    smolClearError();
    i = -5;
  } else if (i <= 0) {
    return same_error(fn);
  }
  LCHECK(state >= 0 && state < kMS, fn, ECbounds, "invalid state");
  LCHECK(state1 >= 0 && state1 < kMS1, fn, ECbounds, "invalid state1");
  LCHECK(state == MSsoln || state1 == MSsoln || state1 == MSbsoln || state1 == state, fn, ECsyntax,
         "nonsensical state combination");
  LCHECK(state2 >= 0 && state2 < kMS1, fn, ECbounds, "invalid state2");
  LCHECK(state1 != state2, fn, ECsyntax, "cannot set rate for state1 = state2");
  int i2 = -5;
  if (newspecies && newspecies[0]) {
    i2 = smolGetSpeciesIndexNT(sim, newspecies);
    LCHECK(i2 > 0, fn, ECerror, "invalid newspecies");
  }
  LCHECK(rate >= 0, fn, ECbounds, "rate needs to be non-negative");
  LCHECK(!(isinternal && rate > 1), fn, ECbounds, "internal rate needs to be <= 1");
  int nsp = (int)sim->spname.size();
  for (int k = 0; k < (int)sim->surf.srf.size(); k++)
    if (s < 0 || k == s) {
      int er = sim->surf.set_rate(k, i, nsp, state, state1, state2, i2, rate, isinternal ? 2 : 1);
      LCHECK(!er, fn, ECerror, "error in surfsetrate");
    }
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolAddPanel(simptr sim, const char* surface, enum PanelShape panelshape, const char* panel,
                            const char* axisstring, double* params) {
  const char* fn = "smolAddPanel";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int s = smolGetSurfaceIndexNT(sim, surface);
  if (s < 0) return same_error(fn);
  LCHECK(panelshape >= 0 && panelshape < PSall, fn, ECnonexist, "invalid panel shape");
  if (panelshape == PSrect) LCHECK(axisstring, fn, ECmissing, "missing axisstring");
  LCHECK(params, fn, ECmissing, "missing params");
  int er = sim->surf.add_panel(s, panelshape, panel, axisstring, params);
  LCHECK(er != -3, fn, ECsyntax, "cannot parse axisstring");
  LCHECK(er != -4, fn, ECbounds, "drawing slices and stacks need to be positive");
  LCHECK(er != -5, fn, ECbounds, "cylinder ends cannot be at the same location");
  LCHECK(er != -6, fn, ECbounds, "hemisphere outward vector has zero length");
  LCHECK(er != -7, fn, ECbounds, "radius needs to be positive");
  LCHECK(er != -8, fn, ECbounds, "normal vector has zero length");
  LCHECK(er != -9, fn, ECerror, "panel name was used before for a different panel shape");
  LCHECK(er >= 0, fn, ECbug, "bug in smolAddPanel");
  invalidate(sim);
  return ECok;
}

int smolGetPanelIndexNT(simptr sim, const char* surface, enum PanelShape* panelshapeptr, const char* panel) {
  if (!sim || !panel) return (int)ECmissing;
  int s = smolGetSurfaceIndexNT(sim, surface);
  if (s < 0) return s;
  if (!strcmp(panel, "all")) return (int)ECall;
  int id = sim->surf.find_panel(s, panel);
  if (id < 0) return (int)ECnonexist;
  const HostPanel& p = sim->surf.pnl[id];
  if (panelshapeptr) *panelshapeptr = (enum PanelShape)p.ps;
  const std::vector<int>& v = sim->surf.srf[s]->panels[p.ps];
  for (size_t k = 0; k < v.size(); k++)
    if (v[k] == id) return (int)k;
  return (int)ECbug;
}

int smolGetPanelIndex(simptr sim, const char* surface, enum PanelShape* panelshapeptr, const char* panel) {
  int p = smolGetPanelIndexNT(sim, surface, panelshapeptr, panel);
  if (p < 0) smolSetError("smolGetPanelIndex", (ErrorCode)p, "panel not found", "");
  return p;
}

enum ErrorCode smolSetPanelJump(simptr sim, const char* surface, const char* panel1, enum PanelFace face1,
                                const char* panel2, enum PanelFace face2, int isbidirectional) {
  const char* fn = "smolSetPanelJump";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int s = smolGetSurfaceIndexNT(sim, surface);
  if (s < 0) return same_error(fn);
  LCHECK(panel1 && panel2, fn, ECmissing, "missing panel name");
  int p1 = sim->surf.find_panel(s, panel1), p2 = sim->surf.find_panel(s, panel2);
  LCHECK(p1 >= 0 && p2 >= 0, fn, ECnonexist, "panel not found");
  LCHECK(face1 == PFfront || face1 == PFback, fn, ECbounds, "face1 must be front or back");
  LCHECK(face2 == PFfront || face2 == PFback, fn, ECbounds, "face2 must be front or back");
  LCHECK(sim->surf.pnl[p1].ps == sim->surf.pnl[p2].ps, fn, ECerror, "jumping panels must have the same shape");
  sim->surf.set_jump(p1, face1, p2, face2, isbidirectional);
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolSetSurfaceSimParams(simptr sim, const char* parameter, double value) {
  const char* fn = "smolSetSurfaceSimParams";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(parameter, fn, ECmissing, "missing parameter name");
  if (!strcmp(parameter, "epsilon")) {
    LCHECK(value > 0, fn, ECbounds, "epsilon needs to be positive");
    sim->surf.epsilon = value;
  } else if (!strcmp(parameter, "margin")) {
    LCHECK(value >= 0, fn, ECbounds, "margin needs to be non-negative");
    sim->surf.margin = value;
  } else if (!strcmp(parameter, "neighbordist")) {
    sim->surf.neighdist = value;
  } else {
    LCHECK(0, fn, ECsyntax, "parameter name not recognized");
  }
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolAddReaction(simptr sim, const char* reaction, const char* reactant1, enum MolecState rstate1,
                               const char* reactant2, enum MolecState rstate2, int nproduct,
                               const char** productspecies, enum MolecState* productstates, double rate) {
  const char* fn = "smolAddReaction";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(reaction && reaction[0], fn, ECmissing, "missing reaction name");
  LCHECK(!sim->started, fn, ECerror, "reactions are fixed once stepping started");
  for (int o = 0; o < 3; o++)
    for (auto& r : sim->rxn[o]) LCHECK(r.name != reaction, fn, ECerror, "reaction name already used");
  Rxn r;
  r.name = reaction;
  int order = 0;
  if (reactant1 && reactant1[0]) {
    int i = smolGetSpeciesIndexNT(sim, reactant1);
    if (i <= 0) return same_error(fn);
    r.rct[order] = i;
    r.rctstate[order] = rstate1;
    order++;
  }
  if (reactant2 && reactant2[0]) {
    int i = smolGetSpeciesIndexNT(sim, reactant2);
    if (i <= 0) return same_error(fn);
    r.rct[order] = i;
    r.rctstate[order] = rstate2;
    order++;
  }
  r.order = order;
  LCHECK(nproduct >= 0 && nproduct <= kMaxProd, fn, ECbounds, "nproduct out of range");
  r.nprod = nproduct;
  for (int p = 0; p < nproduct; p++) {
    int i = smolGetSpeciesIndexNT(sim, productspecies[p]);
    if (i <= 0) return same_error(fn);
    r.prd[p] = i;
    r.prdstate[p] = productstates ? productstates[p] : MSsoln;
  }
  set_permit(r);
  if (rate >= 0) {
    r.rate = rate;
    r.multiplicity = 1;
  }
  int idx = (int)sim->rxn[order].size();
  sim->rxn[order].push_back(r);
  // nrxn/table over distinct permutations of the reactants (RxnAddReaction,
  // smolreact.c:2118-2135)
  if (order >= 1) {
    if (sim->table[order].empty()) sim->table[order].resize(order == 1 ? kMaxSpecies : kMaxSpecies * kMaxSpecies);
    sim->table[order][pack(order, r.rct)].push_back(idx);
    if (order == 2 && r.rct[0] != r.rct[1]) {
      int sw[2] = {r.rct[1], r.rct[0]};
      sim->table[order][pack(order, sw)].push_back(idx);
    }
  }
  invalidate(sim);
  return ECok;
}

// smolGetReactionIndexNT (libsmoldyn.cpp:1846-1867)
int smolGetReactionIndexNT(simptr sim, int* orderptr, const char* reaction) {
  const char* fn = "smolGetReactionIndexNT";
  if (!sim) return set_error_nt(fn, ECmissing, "missing sim");
  if (!reaction) return set_error_nt(fn, ECmissing, "missing reaction");
  if (!strcmp(reaction, "all")) return set_error_nt(fn, ECall, "reaction cannot be 'all'");
  if (orderptr && *orderptr >= 0 && *orderptr < 3) {
    const int o = *orderptr;
    if (sim->rxn[o].empty()) return set_error_nt(fn, ECnonexist, "no reactions defined of this order");
    for (size_t r = 0; r < sim->rxn[o].size(); r++)
      if (sim->rxn[o][r].name == reaction) return (int)r;
    return set_error_nt(fn, ECnonexist, "reaction not found");
  }
  for (int o = 0; o < 3; o++)
    for (size_t r = 0; r < sim->rxn[o].size(); r++)
      if (sim->rxn[o][r].name == reaction) {
        if (orderptr) *orderptr = o;
        return (int)r;
      }
  return set_error_nt(fn, ECnonexist, "reaction not found");
}

int smolGetReactionIndex(simptr sim, int* orderptr, const char* reaction) {
  int r = smolGetReactionIndexNT(sim, orderptr, reaction);
  if (r < 0) return (int)same_error("smolGetReactionIndex");
  return r;
}

enum ErrorCode smolSetReactionRate(simptr sim, const char* reaction, double rate, int type) {
  const char* fn = "smolSetReactionRate";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(!sim->started, fn, ECerror, "rates are fixed once stepping started");
  int order = -1;
  int r = smolGetReactionIndexNT(sim, &order, reaction);
  if (r < 0) return same_error(fn);
  Rxn& rx = sim->rxn[order][r];
  if (type == 0) {
    LCHECK(rate >= 0, fn, ECbounds, "rate must be >=0");
    rx.rate = rate;
    if (rx.multiplicity < 0) rx.multiplicity = 1;
  } else if (type == 1) {
    rx.bindrad2 = rate * rate;
  } else if (type == 2) {
    rx.prob = rate;
  } else {
    LCHECK(0, fn, ECbounds, "type out of range");
  }
  invalidate(sim);
  return ECok;
}

// libsmoldyn.cpp:1990-2020 with RxnSetRevparam (smolreact.c:1854-1880)
enum ErrorCode smolamdSetReactionLog(simptr sim, const char* filename, const char* reaction, const long* serno,
                                     int nserno, int turnon) {
  const char* fn = "smolamdSetReactionLog";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(filename && filename[0], fn, ECmissing, "missing file name");
  LCHECK(reaction && reaction[0], fn, ECmissing, "missing reaction name");
  LCHECK(nserno < 0 || serno, fn, ECmissing, "missing serial numbers");
  bool found = false;
  for (int o = 0; o < 3; o++)
    for (auto& rx : sim->rxn[o]) {
      if (strcmp(reaction, "all") && rx.name != reaction) continue;
      found = true;
      if (turnon) {
        if (nserno < 0) {
          rx.logall = true;
          rx.logserno.clear();
        } else {
          rx.logserno.insert(rx.logserno.end(), serno, serno + nserno);
        }
        rx.logfile = filename;
      } else if (nserno < 0) {
        rx.logall = false;
        rx.logserno.clear();
        rx.logfile.clear();
      } else {
        for (int k = 0; k < nserno; k++)
          rx.logserno.erase(std::remove(rx.logserno.begin(), rx.logserno.end(), serno[k]), rx.logserno.end());
        if (rx.logserno.empty() && !rx.logall) rx.logfile.clear();
      }
    }
  LCHECK(found, fn, ECnonexist, "reaction '%s' not recognized", reaction);
  return ECok;
}

enum ErrorCode smolSetReactionProducts(simptr sim, const char* reaction, enum RevParam method, double parameter,
                                       const char* product, double* position) {
  const char* fn = "smolSetReactionProducts";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int order = -1;
  int r = smolGetReactionIndexNT(sim, &order, reaction);
  if (r < 0) return same_error(fn);
  Rxn& rx = sim->rxn[order][r];
  int prd = -1;
  if (product && product[0]) {
    int i = smolGetSpeciesIndexNT(sim, product);
    if (i <= 0) return same_error(fn);
    for (int p = 0; p < rx.nprod && prd < 0; p++)
      if (rx.prd[p] == i) prd = p;
    LCHECK(prd >= 0, fn, ECerror, "listed product is not a product of the listed reaction");
  }
  int er = rx.rparamt != RPnone ? 1 : 0;
  rx.rparamt = method;
  if (method == RPnone || method == RPirrev || method == RPconfspread) {
  } else if (method == RPbounce) {
    rx.rparam = parameter;
  } else if (method == RPpgem || method == RPpgemmax || method == RPpgemmaxw || method == RPpgem2 ||
             method == RPpgemmax2) {
    if (!(parameter > 0 && parameter <= 1)) er = 2;
    rx.rparam = parameter;
  } else if (method == RPratio || method == RPunbindrad || method == RPratio2) {
    if (parameter < 0) er = 2;
    rx.rparam = parameter;
  } else if (method == RPoffset || method == RPfixed) {
    er = 0;
    if (prd < 0) er = 4;
    else if (!position) er = 5;
    else
      for (int d = 0; d < sim->dim; d++) rx.prdpos[prd][d] = position[d];
  } else {
    er = 3;
  }
  invalidate(sim);
  LCHECK(er != 1, fn, ECwarning, "reaction product parameter was set before");
  LCHECK(er != 2, fn, ECbounds, "reaction product parameter out of bounds");
  LCHECK(er != 3, fn, ECnonexist, "invalid reaction product method");
  LCHECK(er != 4, fn, ECmissing, "missing product name");
  LCHECK(er != 5, fn, ECmissing, "missing product position");
  return g_warncode;
}

// ---------------------------------------------------------------------------
// extensions

enum ErrorCode smolamdAddMoleculesAt(simptr sim, const char* species, int number, const double* positions) {
  const char* fn = "smolamdAddMoleculesAt";
  LCHECK(sim, fn, ECmissing, "missing sim");
  int i = smolGetSpeciesIndexNT(sim, species);
  if (i <= 0) return same_error(fn);
  LCHECK(number >= 0 && (number == 0 || positions), fn, ECmissing, "missing positions");
  if (sim->started) return add_running(sim, fn, i, number, nullptr, nullptr, positions);
  for (int k = 0; k < number; k++) {
    PendingMol pm;
    pm.ident = i;
    pm.serno = sim->serno++;
    for (int d = 0; d < 3; d++) pm.pos[d] = positions[3 * k + d];
    sim->pending.push_back(pm);
  }
  invalidate(sim);
  return ECok;
}

enum ErrorCode smolamdInjectDisplacements(simptr sim, const double* gauss, long nmol) {
  const char* fn = "smolamdInjectDisplacements";
  LCHECK(sim, fn, ECmissing, "missing sim");
  ErrorCode ec;
  int code = run_one(sim, gauss, nmol, fn, &ec);
  if (code < 0) return ec;
  if (code == 1) set_error(fn, ECnotify, "Simulation complete");
  return ECok;
}

// smolamdInjectDisplacements without device allocation (no counterpart in
// the reference; for callers that cannot allocate device memory): the
// simulation must already be set up by smolUpdateSim (its device state exists)
// and at most engine_noalloc_inject_max() molecules may diffuse.
enum ErrorCode smolamdInjectDisplacementsNoAlloc(simptr sim, const double* gauss, long nmol) {
  const char* fn = "smolamdInjectDisplacementsNoAlloc";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(gauss, fn, ECmissing, "missing deviates");
  LCHECK(sim->ok && sim->dev, fn, ECerror,
         "the simulation is not set up: call smolUpdateSim first (this entry point allocates nothing)");
  LCHECK(nmol <= engine_noalloc_inject_max(), fn, ECbounds,
         "%ld molecules exceed the allocation-free limit of %ld diffusing molecules", nmol,
         engine_noalloc_inject_max());
  sim->started = true;
  const int bits = engine_step_noalloc(sim->dev, gauss, nmol);
  if (bits == -4) {
    set_error(fn, ECbounds, "more diffusing molecules than the allocation-free limit of %ld",
              engine_noalloc_inject_max());
    return ECbounds;
  }
  if (bits) return engine_error(fn, bits);
  sim->time += sim->dt;
  sim->stepindex++;
  if (sim->time >= sim->tmax) set_error(fn, ECnotify, "Simulation complete");
  return ECok;
}

int smolamdGetListCount(simptr sim) {
  if (!sim) return (int)ECmissing;
  if (do_update(sim, "smolamdGetListCount") != ECok) return (int)ECerror;
  return (int)sim->listname.size();
}

int smolamdGetMolecules(simptr sim, int list, int* species, double* pos, unsigned long long* serno, int* box,
                        int maxmol) {
  if (!sim) return (int)ECmissing;
  if (do_update(sim, "smolamdGetMolecules") != ECok) return (int)ECerror;
  return engine_get_list(sim->dev, list, species, pos, (uint64_t*)serno, box, maxmol);
}

int smolamdGetReactionLog(simptr sim, long long* rows, int maxrows) {
  if (!sim || !sim->dev) return 0;
  std::vector<LogRow> lr(maxrows > 0 ? maxrows : 1);
  int n;
  if (sim->holdlog) {  // the last step ran on an engine since rebuilt
    n = (int)sim->heldlog.size();
    for (int k = 0; k < n && k < maxrows; k++) lr[(size_t)k] = sim->heldlog[(size_t)k];
  } else {
    n = engine_get_log(sim->dev, lr.data(), maxrows);
  }
  for (int k = 0; k < n && k < maxrows; k++) {
    rows[6 * k + 0] = lr[k].order;
    rows[6 * k + 1] = lr[k].r;
    rows[6 * k + 2] = (long long)lr[k].rct[0];
    rows[6 * k + 3] = (long long)lr[k].rct[1];
    rows[6 * k + 4] = (long long)lr[k].prd[0];
    rows[6 * k + 5] = (long long)lr[k].prd[1];
  }
  return n;
}

enum ErrorCode smolamdGetReactionParams(simptr sim, int order, int reaction, double* out) {
  const char* fn = "smolamdGetReactionParams";
  LCHECK(sim, fn, ECmissing, "missing sim");
  ErrorCode ec = do_update(sim, fn);
  if (ec != ECok) return ec;
  LCHECK(order >= 0 && order <= 2 && reaction >= 0 && reaction < (int)sim->rxn[order].size(), fn, ECbounds,
         "no such reaction");
  Rxn& rx = sim->rxn[order][reaction];
  out[0] = rx.bindrad2;
  out[1] = rx.prob;
  out[2] = rx.unbindrad;
  out[3] = rx.prdpos[0][0];
  out[4] = rx.nprod > 1 ? rx.prdpos[1][0] : 0;
  out[5] = rx.rparamt;
  return ECok;
}

int smolamdGetBoxes(simptr sim, int* side, double* boxmin, double* boxsize) {
  if (!sim) return (int)ECmissing;
  if (do_update(sim, "smolamdGetBoxes") != ECok) return (int)ECerror;
  for (int d = 0; d < 3; d++) {
    side[d] = sim->geo.side[d];
    boxmin[d] = sim->geo.bmin[d];
    boxsize[d] = sim->geo.bsize[d];
  }
  return sim->geo.nbox;
}

enum ErrorCode smolamdRunSteps(simptr sim, long nsteps, int use_graph, double* ms) {
  const char* fn = "smolamdRunSteps";
  LCHECK(sim, fn, ECmissing, "missing sim");
  ErrorCode ec = do_update(sim, fn);
  if (ec != ECok) return ec;
  sim->started = true;
  int bits = engine_run(sim->dev, nsteps, use_graph);
  if (bits) return engine_error(fn, bits);
  for (long k = 0; k < nsteps; k++) sim->time += sim->dt;
  sim->stepindex += (uint32_t)nsteps;
  if (ms) *ms = engine_last_ms(sim->dev);
  return ECok;
}

enum ErrorCode smolamdProfileSteps(simptr sim, long nsteps, double* out) {
  const char* fn = "smolamdProfileSteps";
  LCHECK(sim, fn, ECmissing, "missing sim");
  ErrorCode ec = do_update(sim, fn);
  if (ec != ECok) return ec;
  sim->started = true;
  int bits = engine_profile(sim->dev, nsteps, out);
  if (bits) return engine_error(fn, bits);
  for (long k = 0; k < nsteps; k++) sim->time += sim->dt;
  sim->stepindex += (uint32_t)nsteps;
  return ECok;
}

enum ErrorCode smolamdGetDiagnostics(simptr sim, double* out) {
  const char* fn = "smolamdGetDiagnostics";
  LCHECK(sim, fn, ECmissing, "missing sim");
  LCHECK(sim->dev, fn, ECerror, "simulation not set up");
  Control c;
  LCHECK(engine_sync_control(sim->dev, &c) == 0, fn, ECerror, "HIP runtime error");
  double n = 0;
  for (size_t l = 0; l < sim->listname.size(); l++) n += c.nl[l];
  out[0] = n;
  out[1] = c.maxedges_seen;
  out[2] = c.rounds;
  out[3] = c.err;
  out[4] = (double)c.serno_next;
  out[5] = c.step;
  out[6] = (double)(sim->molsteps_base + c.molsteps);
  out[7] = sim->rebuilds;
  return ECok;
}

double smolamdBindingRadius(double rate, double dt, double difc, double b, int rel) {
  return rxnparam::bindingradius(rate, dt, difc, b, rel);
}
double smolamdNumRxnRate(double step, double a, double b) { return rxnparam::numrxnrate(step, a, b); }
double smolamdUnbindingRadius(double pgem, double dt, double difc, double a) {
  return rxnparam::unbindingradius(pgem, dt, difc, a);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ensembles: many independent simulations stepped together on one device
// (no counterpart in the reference, which runs replicates as separate
// processes)

struct smolamd_ensemble {
  std::vector<simptr> sims;
  DeviceState* dev = nullptr;
  int nlist0 = 0;
  std::vector<std::string> spname;
};

extern "C" ensembleptr smolamdEnsembleNew(simptr* sims, int n) {
  const char* fn = "smolamdEnsembleNew";
  if (!sims || n < 1) {
    set_error(fn, ECmissing, "no simulations");
    return nullptr;
  }
  std::vector<EngineConfig> cfg(n);
  std::vector<DevTables> tab(n);
  std::vector<Geometry> geo(n);
  std::vector<uint32_t> k0(n), k1(n);
  for (int k = 0; k < n; k++) {
    simptr sim = sims[k];
    if (!sim) {
      set_error(fn, ECmissing, "missing simulation %d", k);
      return nullptr;
    }
    if (do_update(sim, fn, true) != ECok) return nullptr;
    if (sim->spname != sims[0]->spname || sim->dt != sims[0]->dt) {
      set_error(fn, ECerror, "ensemble members need the same species and time step (member %d differs)", k);
      return nullptr;
    }
    cfg[k] = sim->prep.cfg;
    tab[k] = sim->prep.tab;
    geo[k] = sim->prep.geo;
    k0[k] = (uint32_t)sim->key;
    k1[k] = (uint32_t)(sim->key >> 32);
  }
  char err[256] = "";
  DeviceState* dev = engine_create_batch(n, cfg.data(), tab.data(), geo.data(), k0.data(), k1.data(),
                                         sims[0]->gtable.data(), 4096, nullptr, err, sizeof(err));
  if (!dev) {
    set_error(fn, ECmemory, "device engine: %s", err);
    return nullptr;
  }
  for (int k = 0; k < n; k++) {
    simptr sim = sims[k];
    for (int l = 0; l < sim->prep.cfg.nlist; l++) {
      int er = engine_upload_list_inst(dev, k, l, (int)sim->prep.ident[l].size(), sim->prep.ident[l].data(),
                                       sim->prep.pos[l].data(), sim->prep.ser[l].data());
      if (er) {
        engine_destroy(dev);
        set_error(fn, ECmemory, "upload of member %d list %d failed", k, l);
        return nullptr;
      }
    }
  }
  for (int k = 0; k < n; k++) engine_set_counters_inst(dev, k, sims[k]->serno, sims[k]->stepindex, sims[k]->counts);
  ensembleptr e = new smolamd_ensemble();
  e->sims.assign(sims, sims + n);
  e->dev = dev;
  e->spname = sims[0]->spname;
  return e;
}

extern "C" void smolamdEnsembleFree(ensembleptr e) {
  if (!e) return;
  engine_destroy(e->dev);
  delete e;
}

extern "C" int smolamdEnsembleSize(ensembleptr e) { return e ? (int)e->sims.size() : 0; }

extern "C" enum ErrorCode smolamdEnsembleRunSteps(ensembleptr e, long nsteps, int use_graph, double* ms) {
  const char* fn = "smolamdEnsembleRunSteps";
  LCHECK(e && e->dev, fn, ECmissing, "missing ensemble");
  int bits = engine_run(e->dev, nsteps, use_graph);
  if (bits) return engine_error(fn, bits);
  if (ms) *ms = engine_last_ms(e->dev);
  for (simptr sim : e->sims) {
    sim->time += sim->dt * (double)nsteps;
    sim->stepindex += (uint32_t)nsteps;
  }
  return ECok;
}

extern "C" int smolamdEnsembleGetMoleculeCount(ensembleptr e, int member, const char* species) {
  const char* fn = "smolamdEnsembleGetMoleculeCount";
  if (!e || member < 0 || member >= (int)e->sims.size()) {
    set_error(fn, ECbounds, "member out of range");
    return (int)ECbounds;
  }
  Control c;
  if (engine_get_control(e->dev, member, &c)) {
    set_error(fn, ECerror, "HIP runtime error");
    return (int)ECerror;
  }
  if (!species || !strcmp(species, "all")) {
    int s = 0;
    for (int i = 1; i < kMaxSpecies; i++) s += c.count[i];
    return s;
  }
  for (size_t i = 1; i < e->spname.size(); i++)
    if (e->spname[i] == species) return c.count[i];
  set_error(fn, ECnonexist, "species not found");
  return (int)ECnonexist;
}

extern "C" int smolamdEnsembleGetMolecules(ensembleptr e, int member, int list, int* species, double* pos,
                                           unsigned long long* serno, int* box, int maxmol) {
  if (!e || member < 0 || member >= (int)e->sims.size()) return (int)ECbounds;
  return engine_get_list_inst(e->dev, member, list, species, pos, (uint64_t*)serno, box, maxmol);
}

extern "C" enum ErrorCode smolamdEnsembleGetDiagnostics(ensembleptr e, int member, double* out) {
  const char* fn = "smolamdEnsembleGetDiagnostics";
  LCHECK(e && member >= 0 && member < (int)e->sims.size(), fn, ECbounds, "member out of range");
  Control c;
  LCHECK(engine_get_control(e->dev, member, &c) == 0, fn, ECerror, "HIP runtime error");
  double n = 0;
  for (int l = 0; l < e->sims[member]->prep.cfg.nlist; l++) n += c.nl[l];
  out[0] = n;
  out[1] = c.maxedges_seen;
  out[2] = c.rounds;
  out[3] = c.err;
  out[4] = (double)c.serno_next;
  out[5] = c.step;
  out[6] = (double)c.molsteps;
  return ECok;
}

namespace smolamd {

int sim_dim(simptr sim) { return sim->dim; }

void sim_bounds(simptr sim, double* lo, double* hi) {
  for (int d = 0; d < 3; d++) {
    lo[d] = sim->low[d];
    hi[d] = sim->high[d];
  }
}

std::vector<std::string> sim_species_names(simptr sim) {
  return std::vector<std::string>(sim->spname.begin() + 1, sim->spname.end());
}

std::string sim_product_name(simptr sim, const char* reaction, int prd) {
  for (int o = 0; o < 3; o++)
    for (auto& r : sim->rxn[o])
      if (r.name == reaction) return prd >= 0 && prd < r.nprod ? sim->spname[(size_t)r.prd[prd]] : std::string();
  return std::string();
}

int sim_reaction_shape(simptr sim, const char* reaction, int* nprod) {
  for (int o = 0; o < 3; o++)
    for (auto& r : sim->rxn[o])
      if (r.name == reaction) {
        if (nprod) *nprod = r.nprod;
        return o;
      }
  return -1;
}

void sim_set_accuracy(simptr sim, double accuracy) {
  sim->accur = accuracy;
  invalidate(sim);
}

void sim_set_output_precision(simptr sim, int precision) { sim->cmds.precision = precision; }

int sim_set_output_format(simptr sim, const char* format) {
  if (!strcmp(format, "ssv") || !strcmp(format, "SSV")) sim->cmds.outformat = 's';
  else if (!strcmp(format, "csv") || !strcmp(format, "CSV")) sim->cmds.outformat = 'c';
  else return 1;
  return 0;
}

int sim_set_file_suffix(simptr sim, const char* name, int suffix) { return cmd::set_suffix(sim->cmds, name, suffix); }

}  // namespace smolamd
