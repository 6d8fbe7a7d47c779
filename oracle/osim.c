/* CPU oracle (test infrastructure only; never linked into the product).
 *
 * Restatement of Smoldyn's per-timestep particle path and the setup it needs.
 * Each function cites the reference function it follows; the data model is in
 * osim.h.  Two random sources are supported:
 *   ORNG_SFMT     the reference's single SFMT stream consumed in program order,
 *                 used to reproduce the reference's own regression outputs
 *                 (examples/S95_regression/output_275) byte for byte;
 *   ORNG_COUNTER  the device engine's counter-based Philox streams
 *                 (include/smoldyn_amd_rng.h), used as the parity oracle for
 *                 the HIP path.
 * Injected Gaussian displacements (osim_inject) replace the diffusion draws of
 * one step in either mode, like swapping sim->diffusefn in the reference
 * (smolsim.cpp:325).
 */
#include "osim.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/smoldyn_amd_rng.h"
#include "oracle.h"

#define O_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* random numbers (random2.h:21-196)                                         */

uint32_t osim_randU32(osim* sim) { return osfmt_u32(&sim->sfmt); }
double osim_randCOD(osim* sim) { return osfmt_u32(&sim->sfmt) * (1.0 / 4294967296.0); }
double osim_randCCD(osim* sim) { return osfmt_u32(&sim->sfmt) * (1.0 / 4294967295.0); }
static double o_randCCD(osim* sim) { return osim_randCCD(sim); }
static double o_randOOD(osim* sim) { return (((double)osfmt_u32(&sim->sfmt)) + 0.5) * (1.0 / 4294967296.0); }

/* ------------------------------------------------------------------------ */
/* setup API                                                                 */

osim* osim_new(int dim, const double* low, const double* high) {
  osim* sim;
  int d, ll, i, ms;
  if (dim < 1 || dim > 3) return NULL;
  sim = (osim*)calloc(1, sizeof(osim));
  if (!sim) return NULL;
  sim->dim = dim;
  for (d = 0; d < dim; d++) {
    sim->wpos[2 * d] = low[d];
    sim->wpos[2 * d + 1] = high[d];
    sim->wtype[2 * d] = sim->wtype[2 * d + 1] = 't'; /* libsmoldyn.cpp:192-195 */
  }
  sim->tmin = sim->time = 0;
  sim->tmax = 100;
  sim->dt = 0.01;
  sim->tbreak = 1e300;
  sim->accur = 10; /* smolsim.cpp:304 */
  sim->nspecies = 1;
  strcpy(sim->spname[0], "empty");
  for (i = 0; i < O_MAXSPECIES; i++)
    for (ms = 0; ms < O_MSMAX; ms++) sim->listlookup[i][ms] = -1;
  for (ll = 0; ll < O_MAXLIST; ll++) sim->listsystem[ll] = 1;
  sim->serno = 1; /* smolmolec.c:1690 */
  sim->mpbox = -1;
  sim->boxsize = -1;
  sim->rngmode = ORNG_SFMT;
  osfmt_seed(&sim->sfmt, 1);
  return sim;
}

static void o_freeboxes(osim* sim) {
  int b, ll;
  if (!sim->boxes) return;
  for (b = 0; b < sim->nbox; b++) {
    free(sim->boxes[b].neigh);
    free(sim->boxes[b].wpneigh);
    for (ll = 0; ll < O_MAXLIST; ll++) free(sim->boxes[b].mol[ll]);
  }
  free(sim->boxes);
  sim->boxes = NULL;
  sim->nbox = 0;
}

void osurf_free(osim* sim);

void osim_free(osim* sim) {
  int ll, m, o, r;
  if (!sim) return;
  for (ll = 0; ll < sim->nlist; ll++) {
    for (m = 0; m < sim->nl[ll]; m++) free(sim->live[ll][m]);
    free(sim->live[ll]);
  }
  for (m = 0; m < sim->nd; m++) free(sim->dead[m]);
  free(sim->dead);
  free(sim->gtable);
  free(sim->gtable_flat);
  o_freeboxes(sim);
  for (o = 0; o < 3; o++)
    for (r = 0; r < sim->rxnss[o].totrxn; r++) free(sim->rxnss[o].rxn[r]);
  osurf_free(sim);
  free(sim->log);
  free(sim);
}

int osim_set_wall(osim* sim, int d, int highside, char type) {
  if (d < 0 || d >= sim->dim) return 1;
  if (!(type == 'r' || type == 'p' || type == 'a' || type == 't')) return 2;
  sim->wtype[2 * d + (highside ? 1 : 0)] = type;
  sim->ok = 0;
  return 0;
}

void osim_set_times(osim* sim, double tstart, double tstop, double dt) {
  sim->tmin = sim->time = tstart;
  sim->tmax = tstop;
  sim->dt = dt;
  sim->ok = 0;
}

void osim_set_seed(osim* sim, long seed) { osfmt_seed(&sim->sfmt, (uint32_t)seed); }

void osim_set_rngmode(osim* sim, int mode, uint64_t key) {
  sim->rngmode = mode;
  sim->key0 = (uint32_t)key;
  sim->key1 = (uint32_t)(key >> 32);
}

int osim_set_partitions(osim* sim, const char* method, double value) {
  if (!strcmp(method, "molperbox")) sim->mpbox = value;
  else if (!strcmp(method, "boxsize")) sim->boxsize = value;
  else return 2;
  sim->ok = 0;
  return 0;
}

void osim_set_accuracy(osim* sim, double accur) { sim->accur = accur; }

int osim_add_species(osim* sim, const char* name) {
  int i = sim->nspecies;
  if (i >= O_MAXSPECIES) return -1;
  strncpy(sim->spname[i], name, O_NAMELEN - 1);
  sim->nspecies++;
  sim->ok = 0;
  return i;
}

int osim_species_index(osim* sim, const char* name) {
  int i;
  for (i = 1; i < sim->nspecies; i++)
    if (!strcmp(sim->spname[i], name)) return i;
  return -1;
}

/* molsetdifc (smolmolec.c:810): ms may be OMS_ALL */
void osim_set_difc(osim* sim, int i, int ms, double difc) {
  int m;
  if (ms == OMS_ALL) {
    for (m = 0; m < O_MSMAX; m++) sim->difc[i][m] = difc;
  } else if (ms >= 0 && ms < O_MSMAX) {
    sim->difc[i][ms] = difc;
  }
  sim->ok = 0;
}

static int o_addlist(osim* sim, const char* name, int system) {
  int ll = sim->nlist;
  if (ll >= O_MAXLIST) return -1;
  strncpy(sim->listname[ll], name, O_NAMELEN - 1);
  sim->listsystem[ll] = system;
  sim->maxl[ll] = 1;
  sim->live[ll] = (omol**)calloc(1, sizeof(omol*));
  sim->nl[ll] = sim->topl[ll] = sim->sortl[ll] = 0;
  sim->nlist++;
  return ll;
}

int osim_add_list(osim* sim, const char* name) { return o_addlist(sim, name, 1); }

int osim_set_list(osim* sim, int i, int ms, int ll) {
  int m;
  if (ll < 0 || ll >= sim->nlist) return 1;
  if (ms == OMS_ALL) {
    for (m = 0; m < O_MSMAX; m++) sim->listlookup[i][m] = ll;
  } else {
    sim->listlookup[i][ms == OMS_BSOLN ? OMS_SOLN : ms] = ll;
  }
  return 0;
}

/* getnextmol (smolmolec.c:2693-2708) with molexpandlist (:1911-1950). */
omol* osim_getnextmol(osim* sim) {
  omol* m;
  if (sim->topd == 0) {
    int nmol = sim->maxd + 1;
    int maxnew = sim->maxd + nmol;
    omol** nl = (omol**)calloc(maxnew, sizeof(omol*));
    int k;
    if (!nl) return NULL;
    for (k = 0; k < sim->maxd; k++) nl[k] = sim->dead[k];
    for (k = sim->nd - 1; k >= sim->topd; k--) {
      nl[k + nmol] = nl[k];
      nl[k] = NULL;
    }
    for (k = sim->topd; k < sim->topd + nmol; k++) {
      nl[k] = (omol*)calloc(1, sizeof(omol));
      nl[k]->list = -1;
      nl[k]->box = -1;
      nl[k]->pnl = nl[k]->pnlx = -1;
    }
    free(sim->dead);
    sim->dead = nl;
    sim->maxd = maxnew;
    sim->topd += nmol;
    sim->nd += nmol;
  }
  m = sim->dead[--sim->topd];
  m->serno = sim->serno++;
  return m;
}

/* molkill (smolmolec.c:2676-2689) */
void osim_molkill(osim* sim, omol* m, int ll, int mi) {
  int d;
  m->ident = 0;
  m->mstate = OMS_SOLN;
  m->list = -1;
  for (d = 0; d < sim->dim; d++) m->posoffset[d] = 0;
  m->pnl = m->pnlx = -1;
  if (ll < 0) {
  } else if (mi < 0) {
    sim->sortl[ll] = 0;
  } else if (mi < sim->sortl[ll]) {
    sim->sortl[ll] = mi;
  }
}

/* pos2box (smolboxes.c:67-79) */
int osim_pos2box(const osim* sim, const double* pos) {
  int b = 0, d, indx;
  for (d = 0; d < sim->dim; d++) {
    indx = (int)((pos[d] - sim->bmin[d]) / sim->bsize[d]);
    if (indx < 0) indx = 0;
    else if (indx >= sim->side[d]) indx = sim->side[d] - 1;
    b = sim->side[d] * b + indx;
  }
  return b;
}

/* addmol (smolmolec.c:2717-2740) with sort 0: positions drawn with unirandOOD
   in every dimension, also where lo==hi (the pointer test at :2726 is never
   true for a 'mol' statement, so fixed coordinates still consume a draw). */
int osim_add_mols(osim* sim, int n, int ident, const double* lo, const double* hi) {
  int k, d;
  for (k = 0; k < n; k++) {
    omol* m = osim_getnextmol(sim);
    if (!m) return 3;
    m->ident = ident;
    m->mstate = OMS_SOLN;
    m->list = sim->listlookup[ident][OMS_SOLN];
    if (sim->rngmode == ORNG_SFMT) {
      for (d = 0; d < sim->dim; d++) m->posx[d] = m->pos[d] = o_randOOD(sim) * (hi[d] - lo[d]) + lo[d];
    } else {
      smolrng4 r = smolrng_mol(sim->key0, sim->key1, m->serno, 0, SMOLRNG_PLACE, 0);
      for (d = 0; d < sim->dim; d++) m->posx[d] = m->pos[d] = smolrng_u01open(r.v[d]) * (hi[d] - lo[d]) + lo[d];
    }
    m->box = sim->nbox ? osim_pos2box(sim, m->pos) : -1;
    m->pnl = m->pnlx = -1;
  }
  /* before stepping: set up with the rest at the next update; once stepping
     started the reference runs no update for them (addmol sorts nothing), so
     they stay resurrected until the next step's first molsort */
  if (sim->stepindex == 0) sim->ok = 0;
  return 0;
}

int osim_add_mols_at(osim* sim, int n, int ident, const double* pos) {
  int k, d;
  for (k = 0; k < n; k++) {
    omol* m = osim_getnextmol(sim);
    if (!m) return 3;
    m->ident = ident;
    m->mstate = OMS_SOLN;
    m->list = sim->listlookup[ident][OMS_SOLN];
    for (d = 0; d < sim->dim; d++) m->posx[d] = m->pos[d] = pos[k * sim->dim + d];
    m->box = sim->nbox ? osim_pos2box(sim, m->pos) : -1;
    m->pnl = m->pnlx = -1;
  }
  /* before stepping: set up with the rest at the next update; once stepping
     started the reference runs no update for them (addmol sorts nothing), so
     they stay resurrected until the next step's first molsort */
  if (sim->stepindex == 0) sim->ok = 0;
  return 0;
}

/* RxnSetPermit (smolreact.c:1899-1922) */
static void o_setpermit(orxn* rxn, int order, const int* rctstate) {
  int ms, nms = order == 1 ? O_MSMAX1 : O_MSMAX1 * O_MSMAX1;
  for (ms = 0; ms < nms; ms++) {
    int msl[2], set = 1, o;
    if (order == 1) msl[0] = ms;
    else {
      msl[0] = ms / O_MSMAX1;
      msl[1] = ms % O_MSMAX1;
    }
    for (o = 0; o < order && set; o++)
      if (!(rctstate[o] == OMS_ALL || rctstate[o] == msl[o])) set = 0;
    if (set) rxn->permit[ms] = 1;
  }
  if (order == 2 && rxn->rct[0] == rxn->rct[1] && rctstate[0] != rctstate[1]) {
    int sw[2];
    sw[0] = rctstate[1];
    sw[1] = rctstate[0];
    for (ms = 0; ms < nms; ms++) {
      int msl0 = ms / O_MSMAX1, msl1 = ms % O_MSMAX1;
      if ((sw[0] == OMS_ALL || sw[0] == msl0) && (sw[1] == OMS_ALL || sw[1] == msl1)) rxn->permit[ms] = 1;
    }
  }
}

/* RxnAddReaction (smolreact.c:2064-2175) + RxnSetValue "rate" (:1746-1750) */
int osim_add_reaction(osim* sim, const char* name, int order, const int* rct, const int* rctstate,
                      int nprod, const int* prd, const int* prdstate, double rate) {
  orxnss* ss;
  orxn* rxn;
  int k, r;
  if (order < 0 || order > 2 || nprod > O_MAXPRD) return -1;
  ss = &sim->rxnss[order];
  ss->present = 1;
  r = ss->totrxn;
  if (r >= O_MAXRXN) return -1;
  rxn = (orxn*)calloc(1, sizeof(orxn));
  strncpy(rxn->name, name, O_NAMELEN - 1);
  rxn->order = order;
  for (k = 0; k < order; k++) {
    rxn->rct[k] = rct[k];
    rxn->rctstate[k] = rctstate[k];
  }
  o_setpermit(rxn, order, rctstate);
  rxn->nprod = nprod;
  for (k = 0; k < nprod; k++) {
    rxn->prd[k] = prd[k];
    rxn->prdstate[k] = prdstate[k];
  }
  rxn->rate = rate;
  rxn->multiplicity = 1;
  rxn->bindrad2 = -1;
  rxn->prob = -1;
  rxn->chi = -1;
  rxn->rparamt = ORP_NONE;
  rxn->unbindrad = -1;
  /* table over the distinct permutations of the reactants (Zn_permute) */
  if (order == 0) {
  } else if (order == 1) {
    int i = rct[0];
    ss->table[i][ss->nrxn[i]++] = r;
  } else {
    int i = rct[0] * O_MAXSPECIES + rct[1];
    ss->table[i][ss->nrxn[i]++] = r;
    if (rct[0] != rct[1]) {
      i = rct[1] * O_MAXSPECIES + rct[0];
      ss->table[i][ss->nrxn[i]++] = r;
    }
  }
  ss->rxn[r] = rxn;
  ss->totrxn++;
  sim->ok = 0;
  return r;
}

/* RxnSetRevparam (smolreact.c:1854-1880) for the modes that use a scalar */
int osim_set_revparam(osim* sim, int order, int r, int rparamt, double rparam) {
  orxn* rxn;
  if (order < 1 || order > 2 || r < 0 || r >= sim->rxnss[order].totrxn) return 1;
  rxn = sim->rxnss[order].rxn[r];
  rxn->rparamt = rparamt;
  if (rparamt != ORP_NONE && rparamt != ORP_IRREV && rparamt != ORP_CONFSPREAD) rxn->rparam = rparam;
  sim->ok = 0;
  return 0;
}

/* RxnSetValue "confspreadrad" (smolreact.c:1762-1765), the confspread_radius
   statement (smolsim.cpp:1736-1755): a conformational-spread reaction whose
   reactants react within the given radius */
int osim_set_confspread(osim* sim, int r, double radius) {
  orxn* rxn;
  if (r < 0 || r >= sim->rxnss[2].totrxn || radius < 0) return 1;
  rxn = sim->rxnss[2].rxn[r];
  rxn->rparamt = ORP_CONFSPREAD;
  rxn->bindrad2 = radius * radius;
  sim->ok = 0;
  return 0;
}

/* RxnSetValue "bindrad" (smolreact.c:1767-1769): a binding radius given
   directly (the reaction's rate left undefined, so rxnsetrate keeps it) */
int osim_set_bindrad(osim* sim, int r, double radius) {
  if (r < 0 || r >= sim->rxnss[2].totrxn || radius < 0) return 1;
  sim->rxnss[2].rxn[r]->bindrad2 = radius * radius;
  sim->ok = 0;
  return 0;
}

/* the position argument of RxnSetRevparam for offset and fixed placement
   (smolreact.c:1870-1874) */
int osim_set_prdpos(osim* sim, int order, int r, int prd, const double* pos) {
  orxn* rxn;
  int d;
  if (order < 1 || order > 2 || r < 0 || r >= sim->rxnss[order].totrxn) return 1;
  rxn = sim->rxnss[order].rxn[r];
  if (prd < 0 || prd >= rxn->nprod || !pos) return 2;
  for (d = 0; d < sim->dim; d++) rxn->prdpos[prd][d] = pos[d];
  sim->ok = 0;
  return 0;
}

/* the counter stream's random rotation, for its statistical tests */
void osim_rng_rotation(uint32_t k0, uint32_t k1, uint64_t serial, uint32_t step, double* m) {
  smolrng_rotation(k0, k1, serial, step, m);
}

int osim_set_rxn_prob(osim* sim, int order, int r, double prob) {
  if (order < 1 || order > 2 || r < 0 || r >= sim->rxnss[order].totrxn) return 1;
  sim->rxnss[order].rxn[r]->prob = prob;
  sim->ok = 0;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* update (simupdate, smolsim.cpp:2367-2473)                                 */

static int o_ismobile(osim* sim, int i, int ms) { return sim->difc[i][ms] > 0; }

/* molsupdatelists (smolmolec.c:2587-2651) and molsupdateparams (:2568-2582) */
static int o_molsupdate(osim* sim) {
  int i, ms, ll;
  if (!sim->gtable) {
    /* molssetgausstable (smolmolec.c:781-805): table, then Fisher-Yates
       shuffle with intrand (random2.c:373-382) */
    int n = 4096, k;
    sim->ngauss = n;
    sim->gtable = (double*)malloc(n * sizeof(double));
    sim->gtable_flat = (double*)malloc(n * sizeof(double));
    omath_randtable(sim->gtable, n);
    memcpy(sim->gtable_flat, sim->gtable, n * sizeof(double));
    if (sim->rngmode == ORNG_SFMT) {
      for (k = n - 1; k > 0; k--) {
        int j = (int)(osim_randU32(sim) % (uint32_t)(k + 1));
        double x = sim->gtable[k];
        sim->gtable[k] = sim->gtable[j];
        sim->gtable[j] = x;
      }
    }
  }
  /* "create system molecule lists if none yet" (smolmolec.c:2614-2628): only
     when no list exists and molecules were ever allocated */
  if (sim->nlist == 0 && sim->maxd > 0 && sim->nspecies > 1) {
    int ndif = 0, nfix = 0;
    for (i = 1; i < sim->nspecies; i++)
      for (ms = 0; ms < O_MSMAX; ms++) {
        if (o_ismobile(sim, i, ms)) ndif = 1;
        else nfix = 1;
      }
    if (ndif) {
      ll = o_addlist(sim, "diffuselist", 1);
      for (i = 1; i < sim->nspecies; i++)
        for (ms = 0; ms < O_MSMAX; ms++)
          if (o_ismobile(sim, i, ms)) sim->listlookup[i][ms] = ll;
    }
    if (nfix) {
      ll = o_addlist(sim, "fixedlist", 1);
      for (i = 1; i < sim->nspecies; i++)
        for (ms = 0; ms < O_MSMAX; ms++)
          if (!o_ismobile(sim, i, ms)) sim->listlookup[i][ms] = ll;
    }
  }
  {
    int ok = 1;
    for (i = 1; i < sim->nspecies && ok; i++)
      for (ms = 0; ms < O_MSMAX && ok; ms++)
        if (sim->listlookup[i][ms] < 0) ok = 0;
    if (!ok) {
      int ull = -1;
      for (ll = 0; ll < sim->nlist; ll++)
        if (!strcmp(sim->listname[ll], "unassignedlist")) ull = ll;
      if (ull < 0) ull = o_addlist(sim, "unassignedlist", 1);
      for (i = 1; i < sim->nspecies; i++)
        for (ms = 0; ms < O_MSMAX; ms++)
          if (sim->listlookup[i][ms] < 0) sim->listlookup[i][ms] = ull;
    }
  }
  {
    int m;
    for (m = sim->topd; m < sim->nd; m++) {
      omol* mp = sim->dead[m];
      mp->list = sim->listlookup[mp->ident][mp->mstate];
    }
  }
  for (ll = 0; ll < sim->nlist; ll++) sim->diffuselist[ll] = 0;
  for (i = 0; i < sim->nspecies; i++)
    for (ms = 0; ms < O_MSMAX; ms++)
      if (o_ismobile(sim, i, ms) && sim->listlookup[i][ms] >= 0) sim->diffuselist[sim->listlookup[i][ms]] = 1;
  for (i = 0; i < sim->nspecies; i++)
    for (ms = 0; ms < O_MSMAX; ms++) sim->difstep[i][ms] = sqrt(2.0 * sim->difc[i][ms] * sim->dt);
  return 0;
}

static int o_molcountall(osim* sim) {
  int ll, c = sim->nd - sim->topd;
  for (ll = 0; ll < sim->nlist; ll++) {
    int m;
    for (m = 0; m < sim->nl[ll]; m++)
      if (sim->live[ll][m]->ident > 0) c++;
  }
  return c;
}

/* neighborZV type 7 (source/libSteve/Zn.c:258-278).  wp[] is the caller's
   wall array read for the first dim entries only, as the reference does. */
static int o_neighbors7(const int* indx, int* c, const int* side, int dim, const int* wpwall, int* wrap,
                        int* mid) {
  int i, n = 0, p = 1, j, k;
  int ofst[O_DIMMAX], wp[O_DIMMAX];
  for (k = 0; k < dim; k++) {
    p *= 3;
    wp[k] = wpwall[k];
  }
  for (i = 0; i < p; i++) {
    int add3;
    wrap[n] = 0;
    if (i == p / 2) {
      *mid = n;
      i++;
    }
    add3 = i;
    for (k = dim - 1; k > 0; k--) {
      ofst[k] = add3 % 3;
      add3 /= 3;
    }
    ofst[0] = add3;
    for (j = 0; j < dim; j++) {
      ofst[j] += indx[j] - 1;
      if (ofst[j] < 0 && wp[j]) {
        ofst[j] = side[j] - 1;
        wrap[n] += 1 << (2 * j);
      } else if (ofst[j] >= side[j] && wp[j]) {
        ofst[j] = 0;
        wrap[n] += 2 << (2 * j);
      } else if (ofst[j] < 0 || ofst[j] >= side[j]) {
        j = dim + 1;
      }
    }
    if (j == dim) {
      int add = ofst[0];
      for (k = 1; k < dim; k++) add = side[k] * add + ofst[k];
      c[n++] = add;
    }
  }
  return n;
}

/* boxesupdatelists (smolboxes.c:717-819) + boxesupdateparams (:634-713) */
static int o_boxesupdate(osim* sim) {
  int dim = sim->dim, d, b, ll, nbox = 1;
  double flt1, mpbox;
  mpbox = sim->mpbox;
  if (mpbox < 0 && sim->boxsize < 0) mpbox = 4; /* boxsetsize(sim,"molperbox",4) */
  if (mpbox < 0) mpbox = 0;
  if (mpbox <= 0 && sim->boxsize <= 0) mpbox = 5;
  if (mpbox > 0) {
    double vol = 1.0, n;
    for (d = 0; d < dim; d++) vol *= sim->wpos[2 * d + 1] - sim->wpos[2 * d];
    n = (double)o_molcountall(sim);
    flt1 = pow(n / mpbox / vol, 1.0 / dim);
  } else {
    flt1 = 1.0 / sim->boxsize;
  }
  o_freeboxes(sim);
  for (d = 0; d < dim; d++) {
    sim->bmin[d] = sim->wpos[2 * d];
    sim->side[d] = (int)ceil((sim->wpos[2 * d + 1] - sim->wpos[2 * d]) * flt1);
    if (!sim->side[d]) sim->side[d] = 1;
    sim->bsize[d] = (sim->wpos[2 * d + 1] - sim->wpos[2 * d]) / sim->side[d];
    nbox *= sim->side[d];
  }
  sim->nbox = nbox;
  sim->boxes = (obox*)calloc(nbox, sizeof(obox));
  for (b = 0; b < nbox; b++) {
    obox* bp = &sim->boxes[b];
    int add = b;
    for (d = dim - 1; d > 0; d--) {
      bp->indx[d] = add % sim->side[d];
      add /= sim->side[d];
    }
    bp->indx[0] = add;
    for (ll = 0; ll < O_MAXLIST; ll++) {
      bp->maxmol[ll] = 1;
      bp->mol[ll] = (omol**)calloc(1, sizeof(omol*));
    }
  }
  if (sim->accur >= 3) {
    int wptemp[27], ntemp[27], w, b2, anyw;
    for (b = 0; b < nbox; b++) {
      obox* bp = &sim->boxes[b];
      int nneigh;
      memset(wptemp, 0, sizeof(wptemp));
      for (w = 0; w < 2 * dim; w++) wptemp[w] = (sim->wtype[w] == 'p' && sim->srfss == NULL);
      if (sim->accur < 9) return 4; /* neighbor types 0 and 6 are not restated */
      {
        int wpwall[O_DIMMAX * 2];
        for (w = 0; w < 2 * dim; w++) wpwall[w] = wptemp[w];
        nneigh = o_neighbors7(bp->indx, ntemp, sim->side, dim, wpwall, wptemp, &bp->midneigh);
      }
      anyw = 0;
      for (b2 = 0; b2 < nneigh; b2++) anyw += wptemp[b2];
      bp->nneigh = nneigh;
      bp->neigh = (int*)calloc(nneigh > 0 ? nneigh : 1, sizeof(int));
      for (b2 = 0; b2 < nneigh; b2++) bp->neigh[b2] = ntemp[b2];
      if (anyw) {
        bp->wpneigh = (int*)calloc(nneigh, sizeof(int));
        for (b2 = 0; b2 < nneigh; b2++) bp->wpneigh[b2] = wptemp[b2];
      }
    }
  }
  /* boxesupdateparams: molecule boxes, then box lists from live lists */
  {
    int ll1, m;
    for (ll1 = -1; ll1 < sim->nlist; ll1++) {
      int mlo, mhi;
      omol** ml;
      if (ll1 == -1) {
        mlo = sim->topd;
        mhi = sim->nd;
        ml = sim->dead;
      } else {
        mlo = 0;
        mhi = sim->nl[ll1];
        ml = sim->live[ll1];
      }
      for (m = mlo; m < mhi; m++)
        if (ml[m]->ident > 0) ml[m]->box = osim_pos2box(sim, ml[m]->pos);
    }
    for (ll1 = 0; ll1 < sim->nlist; ll1++)
      for (m = 0; m < sim->nl[ll1]; m++) {
        omol* mp = sim->live[ll1][m];
        if (mp->ident > 0) {
          int l2 = sim->listlookup[mp->ident][mp->mstate];
          obox* bp = &sim->boxes[mp->box];
          if (bp->nmol[l2] == bp->maxmol[l2]) {
            bp->maxmol[l2] = 2 * bp->maxmol[l2] + 1;
            bp->mol[l2] = (omol**)realloc(bp->mol[l2], bp->maxmol[l2] * sizeof(omol*));
          }
          bp->mol[l2][bp->nmol[l2]++] = mp;
        }
      }
  }
  return 0;
}

static void o_boxaddmol(osim* sim, omol* mp, int ll) {
  obox* bp = &sim->boxes[mp->box];
  if (bp->nmol[ll] == bp->maxmol[ll]) {
    bp->maxmol[ll] = 2 * bp->maxmol[ll] + 1;
    bp->mol[ll] = (omol**)realloc(bp->mol[ll], bp->maxmol[ll] * sizeof(omol*));
  }
  bp->mol[ll][bp->nmol[ll]++] = mp;
}

/* boxremovemol (smolboxes.c:167-175): swap with the last entry */
static void o_boxremovemol(osim* sim, omol* mp, int ll) {
  obox* bp = &sim->boxes[mp->box];
  int m;
  for (m = bp->nmol[ll] - 1; m >= 0 && bp->mol[ll][m] != mp; m--)
    ;
  if (m >= 0) bp->mol[ll][m] = bp->mol[ll][--bp->nmol[ll]];
  mp->box = -1;
}

static void o_liveappend(osim* sim, int ll, omol* mp) {
  if (sim->nl[ll] == sim->maxl[ll]) {
    sim->maxl[ll] = 2 * sim->maxl[ll] + 1;
    sim->live[ll] = (omol**)realloc(sim->live[ll], sim->maxl[ll] * sizeof(omol*));
  }
  sim->live[ll][sim->nl[ll]++] = mp;
}

void osurf_removemol(osim* sim, omol* mp, int ll);
int osurf_addmol(osim* sim, omol* mp, int ll);

/* molsort (smolmolec.c:2943-3023) */
static int o_molsort(osim* sim, int onlydead2live) {
  int ll, m, ll2;
  omol** dead;
  if (!onlydead2live) {
    for (ll = 0; ll < sim->nlist; ll++) sim->topl[ll] = sim->nl[ll];
    for (ll = 0; ll < sim->nlist; ll++) {
      omol** ml = sim->live[ll];
      for (m = sim->sortl[ll]; m < sim->topl[ll]; m++) {
        if (ml[m]->list != ll) {
          omol* mp = ml[m];
          if (mp->list == -1) {
            if (mp->box >= 0) o_boxremovemol(sim, mp, ll);
            if (mp->pnl >= 0) osurf_removemol(sim, mp, ll);
            sim->dead[sim->nd++] = sim->dead[sim->topd];
            sim->dead[sim->topd++] = mp;
            ml[m] = NULL;
          } else {
            int bptr = mp->box, pnl = mp->pnl;
            ll2 = mp->list;
            if (bptr >= 0) o_boxremovemol(sim, mp, ll);
            if (pnl >= 0) osurf_removemol(sim, mp, ll);
            o_liveappend(sim, ll2, mp);
            ml = sim->live[ll];
            ml[m] = NULL;
            if (sim->listsystem[ll2]) {
              mp->box = bptr >= 0 ? bptr : osim_pos2box(sim, mp->pos);
              o_boxaddmol(sim, mp, ll2);
              if (pnl >= 0) {
                mp->pnl = pnl;
                osurf_addmol(sim, mp, ll2);
              }
            }
          }
          ml[m] = ml[--sim->topl[ll]];
          ml[sim->topl[ll]] = ml[--sim->nl[ll]];
          ml[sim->nl[ll]] = NULL;
          m--;
        }
      }
    }
  }
  dead = sim->dead;
  for (m = sim->topd; m < sim->nd; m++) {
    omol* mp = dead[m];
    if (mp->ident == 0) {
      dead[sim->topd++] = mp;
    } else {
      ll2 = mp->list;
      o_liveappend(sim, ll2, mp);
      dead[m] = NULL;
      if (sim->listsystem[ll2]) {
        o_boxaddmol(sim, mp, ll2);
      }
    }
  }
  sim->nd = sim->topd;
  if (!onlydead2live)
    for (ll = 0; ll < sim->nlist; ll++) sim->sortl[ll] = sim->nl[ll];
  return 0;
}

/* reassignmolecs (smolboxes.c:896-954) */
static int o_reassign(osim* sim, int diffusing, int reborn) {
  int ll, b, m;
  for (ll = 0; ll < sim->nlist; ll++) {
    if (!sim->listsystem[ll]) continue;
    if (!(diffusing == 0 || sim->diffuselist[ll] == 1)) continue;
    if (!reborn) {
      for (b = 0; b < sim->nbox; b++) sim->boxes[b].nmol[ll] = 0;
      if (sim->srfss) {
        extern void osurf_clearlist(osim * sim, int ll);
        osurf_clearlist(sim, ll);
      }
      for (m = 0; m < sim->nl[ll]; m++) {
        omol* mp = sim->live[ll][m];
        mp->box = osim_pos2box(sim, mp->pos);
        o_boxaddmol(sim, mp, ll);
        if (mp->pnl >= 0) osurf_addmol(sim, mp, ll);
      }
    } else {
      for (m = sim->topl[ll]; m < sim->nl[ll]; m++) {
        omol* mp = sim->live[ll][m];
        int b1 = osim_pos2box(sim, mp->pos);
        if (mp->box != b1) {
          obox* bp = &sim->boxes[mp->box];
          int m2;
          for (m2 = 0; bp->mol[ll][m2] != mp; m2++)
            ;
          bp->mol[ll][m2] = bp->mol[ll][--bp->nmol[ll]];
          mp->box = b1;
          o_boxaddmol(sim, mp, ll);
        }
        if (mp->pnl >= 0) osurf_addmol(sim, mp, ll);
      }
    }
  }
  return 0;
}

static void o_findreverse(osim* sim, int order, int r, int* optr, int* rptr, int* revptr);

/* rxnreactantstate (smolreact.c:277-343), order 2, convertb2f=1 */
static int o_rctstate2(orxn* rxn, int* ms1, int* ms2) {
  const int* p = rxn->permit;
  int a, b, mp;
  if (p[OMS_SOLN * O_MSMAX1 + OMS_SOLN]) { *ms1 = OMS_SOLN; *ms2 = OMS_SOLN; return 1; }
  if (p[OMS_SOLN * O_MSMAX1 + OMS_BSOLN]) { *ms1 = OMS_SOLN; *ms2 = OMS_SOLN; return 1; }
  if (p[OMS_BSOLN * O_MSMAX1 + OMS_SOLN]) { *ms1 = OMS_SOLN; *ms2 = OMS_SOLN; return 1; }
  if (p[OMS_BSOLN * O_MSMAX1 + OMS_BSOLN]) { *ms1 = OMS_SOLN; *ms2 = OMS_SOLN; return 1; }
  for (a = 0; a < O_MSMAX1 && !p[a * O_MSMAX1 + OMS_SOLN]; a++)
    ;
  if (a < O_MSMAX1) { *ms1 = a == OMS_BSOLN ? OMS_SOLN : a; *ms2 = OMS_SOLN; return 1; }
  for (b = 0; b < O_MSMAX1 && !p[OMS_SOLN * O_MSMAX1 + b]; b++)
    ;
  if (b < O_MSMAX1) { *ms1 = OMS_SOLN; *ms2 = b == OMS_BSOLN ? OMS_SOLN : b; return 1; }
  for (mp = 0; mp < O_MSMAX1 * O_MSMAX1 && !p[mp]; mp++)
    ;
  if (mp < O_MSMAX1 * O_MSMAX1) {
    *ms1 = mp / O_MSMAX1;
    *ms2 = mp % O_MSMAX1;
    if (*ms1 == OMS_BSOLN) *ms1 = OMS_SOLN;
    if (*ms2 == OMS_BSOLN) *ms2 = OMS_SOLN;
    return 1;
  }
  *ms1 = *ms2 = OMS_NONE;
  return 0;
}

static double o_difcsum(osim* sim, int i1, int ms1, int i2, int ms2) {
  double sum = 0;
  if (i1) {
    if (ms1 >= O_MSMAX) ms1 = OMS_SOLN;
    sum += sim->difc[i1][ms1];
  }
  if (i2) {
    if (ms2 >= O_MSMAX) ms2 = OMS_SOLN;
    sum += sim->difc[i2][ms2];
  }
  return sum;
}

/* rxnsetrate (smolreact.c:1306-1412), orders 1 and 2 */
static int o_rxnsetrate(osim* sim, int order, int r) {
  orxnss* ss = &sim->rxnss[order];
  orxn* rxn = ss->rxn[r];
  int o2, r2, rev;
  if (rxn->disable) {
    rxn->prob = 0;
  } else if (rxn->rparamt == ORP_CONFSPREAD) {
    if (rxn->rate < 0) return 1;
    rxn->prob = 1.0 - exp(-sim->dt * rxn->rate * rxn->multiplicity);
  } else if (order == 0) {
    /* expected events per step in the system volume (systemvolume, smolwall.c:50-57) */
    double vol = 1.0;
    int d;
    if (rxn->rate < 0) return 1;
    for (d = 0; d < sim->dim; d++) vol *= sim->wpos[2 * d + 1] - sim->wpos[2 * d];
    rxn->prob = rxn->rate * rxn->multiplicity * sim->dt * vol;
  } else if (order == 1) {
    int ms, j, i;
    double sum, product, prob;
    for (ms = 0; ms < O_MSMAX && !rxn->permit[ms]; ms++)
      ;
    if (rxn->rate < 0) return 1;
    if (rxn->rate == 0) rxn->prob = 0;
    else if (ms == O_MSMAX) rxn->prob = 0;
    else {
      i = rxn->rct[0];
      sum = 0;
      for (j = 0; j < ss->nrxn[i]; j++) {
        orxn* rx2 = ss->rxn[ss->table[i][j]];
        if (rx2->permit[ms] && rx2->rate > 0) sum += rx2->rate * rx2->multiplicity;
      }
      rxn->prob = rxn->rate * rxn->multiplicity / sum * (1.0 - exp(-sum * sim->dt));
      product = 1.0;
      for (j = 0; j < ss->nrxn[i]; j++) {
        orxn* rx2 = ss->rxn[ss->table[i][j]];
        if (rx2 == rxn) break;
        if (rx2->permit[ms]) {
          prob = rx2->prob * product;
          product *= 1.0 - prob;
        }
      }
      rxn->prob /= product;
      if (!(rxn->prob >= 0 && rxn->prob <= 1)) return 5;
    }
    o_findreverse(sim, 1, r, &o2, &r2, &rev);
    if (rev > 0 && o2 == 2 && rxn->rparamt == ORP_NONE) {
      rxn->rparamt = ORP_PGEMMAXW;
      rxn->rparam = 0.2;
    }
  } else {
    int i1, i2, ms1, ms2, permit, rparamt;
    double rate3, dsum, rparam, unbindrad;
    if (rxn->rate < 0) {
      if (rxn->prob < 0) rxn->prob = 1;
      return 1;
    }
    i1 = rxn->rct[0];
    i2 = rxn->rct[1];
    permit = o_rctstate2(rxn, &ms1, &ms2);
    rate3 = rxn->rate * rxn->multiplicity;
    if (i1 == i2) rate3 *= 2;
    if ((ms1 == OMS_SOLN && ms2 != OMS_SOLN) || (ms1 != OMS_SOLN && ms2 == OMS_SOLN)) rate3 *= 2;
    dsum = o_difcsum(sim, i1, ms1, i2, ms2);
    o_findreverse(sim, 2, r, &o2, &r2, &rev);
    if (rev > 0 && o2 == 2 && rxn->rparamt == ORP_NONE) {
      rxn->rparamt = ORP_PGEMMAXW;
      rxn->rparam = 0.2;
    }
    if (rev == 1) {
      orxn* revrxn = sim->rxnss[o2].rxn[r2];
      if (revrxn->rparamt == ORP_NONE) {
        revrxn->rparamt = ORP_PGEMMAXW;
        revrxn->rparam = 0.2;
      }
      rparamt = revrxn->rparamt;
      rparam = revrxn->rparam;
    } else {
      rparamt = ORP_NONE;
      rparam = 0;
    }
    if (rxn->prob < 0) rxn->prob = 1;
    if (!permit) rxn->bindrad2 = 0;
    else if (rate3 <= 0) rxn->bindrad2 = 0;
    else if (dsum <= 0) return 4;
    else if (rparamt == ORP_UNBINDRAD)
      rxn->bindrad2 = orxn_bindingradiusprob(rate3, sim->dt, dsum, rparam, 0, rxn->chi, &rxn->prob);
    else if (rparamt == ORP_RATIO)
      rxn->bindrad2 = orxn_bindingradiusprob(rate3, sim->dt, dsum, rparam, 1, rxn->chi, &rxn->prob);
    else if (rparamt == ORP_PGEM)
      rxn->bindrad2 = orxn_bindingradiusprob(rate3 * (1.0 - rparam), sim->dt, dsum, -1, 0, rxn->chi, &rxn->prob);
    else if (rparamt == ORP_PGEMMAX || rparamt == ORP_PGEMMAXW) {
      rxn->bindrad2 = orxn_bindingradiusprob(rate3, sim->dt, dsum, 0, 0, rxn->chi, &rxn->prob);
      unbindrad = orxn_unbindingradius(rparam, sim->dt, dsum, rxn->bindrad2);
      if (unbindrad > 0)
        rxn->bindrad2 = orxn_bindingradiusprob(rate3 * (1.0 - rparam), sim->dt, dsum, -1, 0, rxn->chi, &rxn->prob);
    } else
      rxn->bindrad2 = orxn_bindingradiusprob(rate3, sim->dt, dsum, -1, 0, rxn->chi, &rxn->prob);
    rxn->bindrad2 *= rxn->bindrad2;
  }
  return 0;
}

/* findreverserxn (smolreact.c:345-381) */
static void o_findreverse(osim* sim, int order, int r, int* optr, int* rptr, int* revptr) {
  orxn* rxn = sim->rxnss[order].rxn[r];
  int orderr = 0, rrret = 0, rev = 0, j;
  if (rxn->nprod == 0 || rxn->nprod >= 3 || !sim->rxnss[rxn->nprod].present) {
    rev = 0;
  } else {
    orxnss* ssr;
    int identr, mstater;
    orderr = rxn->nprod;
    ssr = &sim->rxnss[orderr];
    identr = orderr == 1 ? rxn->prd[0] : rxn->prd[0] * O_MAXSPECIES + rxn->prd[1];
    mstater = orderr == 1 ? rxn->prdstate[0] : rxn->prdstate[0] * O_MSMAX1 + rxn->prdstate[1];
    for (j = 0; j < ssr->nrxn[identr]; j++) {
      int rr = ssr->table[identr][j];
      orxn* rxnr = ssr->rxn[rr];
      if (rxnr->permit[mstater]) {
        if (rev != 1 && rxnr->nprod == order) {
          int same;
          if (order == 1) same = rxn->rct[0] == rxnr->prd[0];
          else same = (rxn->rct[0] == rxnr->prd[0] && rxn->rct[1] == rxnr->prd[1]) ||
                      (rxn->rct[0] == rxnr->prd[1] && rxn->rct[1] == rxnr->prd[0]);
          if (same) {
            int identrprd = order == 1 ? rxnr->prd[0] : rxnr->prd[0] * O_MAXSPECIES + rxnr->prd[1];
            int mstaterprd = order == 1 ? rxnr->prdstate[0] : rxnr->prdstate[0] * O_MSMAX1 + rxnr->prdstate[1];
            orxnss* ss = &sim->rxnss[order];
            int jr;
            for (jr = 0; jr < ss->nrxn[identrprd]; jr++)
              if (ss->table[identrprd][jr] == r && ss->rxn[r]->permit[mstaterprd]) {
                rev = 1;
                rrret = rr;
              }
          }
        }
        if (!rev) {
          rev = 2;
          rrret = rr;
        }
      }
    }
  }
  *optr = orderr;
  *rptr = rrret;
  *revptr = rev;
}

/* rxnsetproduct (smolreact.c:1431-1575), non-bounce branches */
static int o_rxnsetproduct(osim* sim, int order, int r) {
  orxn* rxn = sim->rxnss[order].rxn[r];
  int nprod = rxn->nprod, rparamt = rxn->rparamt, d, prd;
  double rpar = rxn->rparam;
  int dim = sim->dim;
  if (nprod == 0) {
    if (rparamt == ORP_NONE || rparamt == ORP_IRREV || rparamt == ORP_CONFSPREAD) rxn->unbindrad = 0;
    else return 1;
  } else if (rparamt == ORP_OFFSET || rparamt == ORP_FIXED) {
    double dist = 0;
    if (nprod == 1)
      for (d = 0; d < dim; d++) dist = rxn->prdpos[0][d];
    else
      for (d = 0; d < dim; d++) dist += rxn->prdpos[0][d] - rxn->prdpos[1][d];
    rxn->unbindrad = sqrt(dist);
  } else if (nprod == 1) {
    if (rparamt == ORP_NONE || rparamt == ORP_IRREV || rparamt == ORP_CONFSPREAD) {
      rxn->unbindrad = 0;
      for (d = 0; d < dim; d++) rxn->prdpos[0][d] = 0;
    } else
      return 2;
  } else if (rparamt == ORP_BOUNCE) {
    /* nprod >= 2, bounce (smolreact.c:1470-1494): offsets from the reactants'
       relative diffusion coefficients (the reflection method, rparam -2,
       takes the rparam < 0 offsets for its fallback) */
    double dc1, dc2, dsum;
    int ms1 = rxn->rctstate[0], ms2 = rxn->rctstate[1];
    if (ms1 == OMS_BSOLN || ms1 == OMS_NONE || ms1 == OMS_ALL || ms1 == OMS_SOME) ms1 = OMS_SOLN;
    if (ms2 == OMS_BSOLN || ms2 == OMS_NONE || ms2 == OMS_ALL || ms2 == OMS_SOME) ms2 = OMS_SOLN;
    dc1 = o_difcsum(sim, rxn->rct[0], ms1, 0, OMS_NONE);
    dc2 = o_difcsum(sim, rxn->rct[1], ms2, 0, OMS_NONE);
    dsum = dc1 + dc2;
    if (dsum == 0) dc1 = dc2 = 0.5;
    else {
      dc1 /= dsum;
      dc2 /= dsum;
    }
    if (rpar >= 0) {
      rxn->unbindrad = rpar;
      rxn->prdpos[0][0] = rpar * dc1;
      rxn->prdpos[1][0] = rpar * dc2;
    } else {
      rxn->unbindrad = rpar;
      rxn->prdpos[0][0] = sqrt(rxn->bindrad2) * dc1;
      rxn->prdpos[1][0] = sqrt(rxn->bindrad2) * dc2;
    }
    for (prd = 2; prd < nprod; prd++) rxn->prdpos[prd][0] = 0;
  } else {
    int ms1 = rxn->prdstate[0], ms2 = rxn->prdstate[1], orderr, rr, rev;
    double dc1, dc2, dsum;
    if (ms1 == OMS_BSOLN) ms1 = OMS_SOLN;
    if (ms2 == OMS_BSOLN) ms2 = OMS_SOLN;
    dc1 = o_difcsum(sim, rxn->prd[0], ms1, 0, OMS_NONE);
    dc2 = o_difcsum(sim, rxn->prd[1], ms2, 0, OMS_NONE);
    dsum = dc1 + dc2;
    if (dsum == 0) dc1 = dc2 = 0.5;
    else {
      dc1 /= dsum;
      dc2 /= dsum;
    }
    o_findreverse(sim, order, r, &orderr, &rr, &rev);
    if (rparamt == ORP_IRREV || rparamt == ORP_CONFSPREAD) {
      rxn->unbindrad = 0;
      for (prd = 0; prd < nprod; prd++)
        for (d = 0; d < dim; d++) rxn->prdpos[prd][d] = 0;
    } else if (rev == 0) {
      if (rparamt == ORP_PGEM || rparamt == ORP_PGEMMAX || rparamt == ORP_PGEMMAXW || rparamt == ORP_RATIO ||
          rparamt == ORP_PGEM2 || rparamt == ORP_PGEMMAX2 || rparamt == ORP_RATIO2)
        return 3;
      else if (rparamt == ORP_UNBINDRAD) {
        rxn->unbindrad = rpar;
        rxn->prdpos[0][0] = rpar * dc1;
        rxn->prdpos[1][0] = -rpar * dc2;
      } else {
        rxn->unbindrad = 0;
        rxn->prdpos[0][0] = 0;
        rxn->prdpos[1][0] = 0;
      }
    } else {
      orxn* rxnr = sim->rxnss[orderr].rxn[rr];
      double bindradr = rxnr->bindrad2 >= 0 ? sqrt(rxnr->bindrad2) : -1;
      if (rparamt == ORP_NONE) return 5;
      else if (rparamt == ORP_UNBINDRAD) {
        rxn->unbindrad = rpar;
        rxn->prdpos[0][0] = rpar * dc1;
        rxn->prdpos[1][0] = -rpar * dc2;
      } else if (rxnr->bindrad2 < 0) return 6;
      else if (rxnr->bindrad2 == 0) {
        rxn->unbindrad = 0;
        rxn->prdpos[0][0] = 0;
        rxn->prdpos[1][0] = 0;
      } else if (rparamt == ORP_RATIO || rparamt == ORP_RATIO2) {
        rxn->unbindrad = rpar * bindradr;
        rxn->prdpos[0][0] = rpar * bindradr * dc1;
        rxn->prdpos[1][0] = -rpar * bindradr * dc2;
      } else if (dsum <= 0) return 4;
      else if (rparamt == ORP_PGEM || rparamt == ORP_PGEM2) {
        rpar = orxn_unbindingradius(rpar, sim->dt, dsum, bindradr);
        if (rpar < 0) return 7;
        rxn->unbindrad = rpar;
        rxn->prdpos[0][0] = rpar * dc1;
        rxn->prdpos[1][0] = -rpar * dc2;
      } else if (rparamt == ORP_PGEMMAX || rparamt == ORP_PGEMMAXW || rparamt == ORP_PGEMMAX2) {
        rpar = orxn_unbindingradius(rpar, sim->dt, dsum, bindradr);
        if (rpar == -2) return 9;
        else if (rpar <= 0) rxn->unbindrad = 0;
        else {
          rxn->unbindrad = rpar;
          rxn->prdpos[0][0] = rpar * dc1;
          rxn->prdpos[1][0] = -rpar * dc2;
        }
      } else
        return 10;
    }
  }
  return 0;
}

/* rxnsupdatelists (smolreact.c:2320-2381) */
static void o_rxnlists(osim* sim, int order) {
  orxnss* ss = &sim->rxnss[order];
  int r, ms1, ms2;
  memset(ss->rxnmollist, 0, sizeof(ss->rxnmollist));
  for (r = 0; r < ss->totrxn; r++) {
    orxn* rxn = ss->rxn[r];
    int i1 = rxn->rct[0];
    if (order == 1) {
      for (ms1 = 0; ms1 < O_MSMAX1; ms1++)
        if (rxn->permit[ms1] && (rxn->prob > 0 || rxn->rate > 0)) {
          int ll1 = sim->listlookup[i1][ms1 == OMS_BSOLN ? OMS_SOLN : ms1];
          ss->rxnmollist[ll1] = 1;
        }
    } else {
      int i2 = rxn->rct[1];
      for (ms1 = 0; ms1 < O_MSMAX1; ms1++)
        for (ms2 = 0; ms2 < O_MSMAX1; ms2++)
          if (rxn->permit[ms1 * O_MSMAX1 + ms2] && rxn->prob != 0 && (rxn->rate > 0 || rxn->bindrad2 > 0)) {
            int ll1 = sim->listlookup[i1][ms1 == OMS_BSOLN ? OMS_SOLN : ms1];
            int ll2 = sim->listlookup[i2][ms2 == OMS_BSOLN ? OMS_SOLN : ms2];
            ss->rxnmollist[ll1 * O_MAXLIST + ll2] = 1;
            ss->rxnmollist[ll2 * O_MAXLIST + ll1] = 1;
          }
    }
  }
}

int osurf_update(osim* sim);

int osim_update(osim* sim) {
  int order, r, er;
  if (sim->ok) return 0;
  er = o_molsupdate(sim);
  if (er) return er;
  er = o_boxesupdate(sim);
  if (er) return 10 + er;
  o_molsort(sim, 0);
  for (order = 1; order <= 2; order++)
    if (sim->rxnss[order].present) o_rxnlists(sim, order);
  for (order = 0; order <= 2; order++)
    for (r = 0; r < sim->rxnss[order].totrxn; r++) {
      er = o_rxnsetrate(sim, order, r);
      if (er > 1) return 100 + er;
    }
  for (order = 0; order <= 2; order++)
    for (r = 0; r < sim->rxnss[order].totrxn; r++) {
      if (order == 0 && sim->rxnss[0].rxn[r]->nprod > 1) return 213; /* one product per zeroth-order reaction */
      er = o_rxnsetproduct(sim, order, r);
      if (er) return 200 + er;
      if (order == 0 && sim->rxnss[0].rxn[r]->nprod == 1) {
        int d;
        for (d = 0; d < sim->dim; d++)
          if (sim->rxnss[0].rxn[r]->prdpos[0][d] != 0) return 214; /* no reactant to key an offset draw */
      }
    }
  if (sim->srfss) {
    er = osurf_update(sim);
    if (er) return 300 + er;
  }
  o_reassign(sim, 0, 0);
  sim->ok = 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* one time step                                                             */

int osim_inject(osim* sim, const double* g, long n) {
  sim->inject = g;
  sim->ninject = n;
  return 0;
}

/* diffuse (smolmolec.c:3115-3168), isotropic solution molecules */
static int o_diffuse(osim* sim) {
  int ll, m, d, dim = sim->dim, nm1 = sim->ngauss - 1;
  long k = 0;
  for (ll = 0; ll < sim->nlist; ll++) {
    if (!sim->diffuselist[ll]) continue;
    for (m = 0; m < sim->nl[ll]; m++) {
      omol* mp = sim->live[ll][m];
      double ds = sim->difstep[mp->ident][mp->mstate];
      for (d = 0; d < dim; d++) mp->posx[d] = mp->pos[d];
      mp->pnlx = mp->pnl;
      if (sim->inject) {
        if (k >= sim->ninject) return 1;
        for (d = 0; d < dim; d++) mp->pos[d] += ds * sim->inject[k * dim + d];
        k++;
      } else if (sim->rngmode == ORNG_SFMT) {
        for (d = 0; d < dim; d++) mp->pos[d] += ds * sim->gtable[osim_randU32(sim) & nm1];
      } else {
        smolrng4 r = smolrng_mol(sim->key0, sim->key1, mp->serno, sim->stepindex, SMOLRNG_DIFFUSE, 0);
        for (d = 0; d < dim; d++) mp->pos[d] += ds * sim->gtable_flat[r.v[d] & (uint32_t)nm1];
      }
      if (mp->mstate != OMS_SOLN) {
        extern void osurf_movemol2closepanel(osim * sim, omol * mp);
        osurf_movemol2closepanel(sim, mp);
      }
    }
  }
  return 0;
}

/* checkwalls (smolwall.c:368-424).  The reference resets m to 0 in every wall
   loop, so the "reborn" call also covers the whole list. */
static void o_checkwalls(osim* sim, int ll, int reborn) {
  int w, m, nmol = sim->nl[ll];
  omol** ml = sim->live[ll];
  for (w = 0; w < 2 * sim->dim; w++) {
    int d = w / 2, side = w % 2;
    double wp = sim->wpos[w];
    char t = sim->wtype[w];
    if (t == 'r' && side == 0) {
      double p2 = 2 * wp;
      for (m = 0; m < nmol; m++)
        if (ml[m]->pos[d] < wp) ml[m]->pos[d] = p2 - ml[m]->pos[d];
    } else if (t == 'r') {
      double p2 = 2 * wp;
      for (m = 0; m < nmol; m++)
        if (ml[m]->pos[d] > wp) ml[m]->pos[d] = p2 - ml[m]->pos[d];
    } else if (t == 'p' && side == 0) {
      double p2 = sim->wpos[w + 1] - wp;
      for (m = 0; m < nmol; m++)
        if (ml[m]->pos[d] < wp) {
          ml[m]->pos[d] += p2;
          ml[m]->posoffset[d] -= p2;
        }
    } else if (t == 'p') {
      double p2 = sim->wpos[w - 1] - wp;
      for (m = 0; m < nmol; m++)
        if (ml[m]->pos[d] > wp) {
          ml[m]->pos[d] += p2;
          ml[m]->posoffset[d] -= p2;
        }
    } else if (t == 'a') {
      for (m = 0; m < nmol; m++) {
        omol* mp = ml[m];
        double diff = wp - mp->pos[d];
        double difi = wp - mp->posx[d];
        double step = sim->difstep[mp->ident][OMS_SOLN];
        int kill;
        if ((!side && diff > 0) || (side && diff < 0)) kill = 1;
        else if (sim->rngmode == ORNG_SFMT) kill = osim_randCOD(sim) < exp(-2 * difi * diff / step / step);
        else {
          smolrng4 r = smolrng_mol(sim->key0, sim->key1, mp->serno, sim->stepindex, SMOLRNG_WALL,
                                   (uint32_t)(w + (reborn ? 8 : 0)));
          kill = smolrng_u01(r.v[0]) < exp(-2 * difi * diff / step / step);
        }
        if (kill) osim_molkill(sim, mp, ll, -1);
      }
    }
  }
}

static void o_logevent(osim* sim, int order, int r, omol* m1, omol* m2, omol** prods, int nprod,
                       const double* rxnpos) {
  orxnevent* e;
  int d;
  if (sim->nlog == sim->maxlog) {
    sim->maxlog = 2 * sim->maxlog + 16;
    sim->log = (orxnevent*)realloc(sim->log, sim->maxlog * sizeof(orxnevent));
  }
  e = &sim->log[sim->nlog++];
  e->order = order;
  e->r = r;
  e->rct[0] = m1 ? m1->serno : 0;
  e->rct[1] = m2 ? m2->serno : 0;
  e->prd[0] = nprod > 0 ? prods[0]->serno : 0;
  e->prd[1] = nprod > 1 ? prods[1]->serno : 0;
  for (d = 0; d < O_DIMMAX; d++) e->pos[d] = d < sim->dim ? rxnpos[d] : 0;
}

/* doreact (smolreact.c:2662-2901), default product-placement branch, no
   surfaces.  m1 is the reactant whose identity is rctident[0]. */
static int o_doreact(osim* sim, int order, int r, omol* m1, omol* m2, int ll1, int mi1, int ll2, int mi2,
                     const double* pos0) {
  orxn* rxn = sim->rxnss[order].rxn[r];
  int dim = sim->dim, d, prd, calc = 0, rxnbox, offaxis = 0;
  double rxnpos[O_DIMMAX], m3[9], v1[O_DIMMAX], u[O_DIMMAX];
  double bv[O_DIMMAX] = {0}, bdist = 0, bx = 0; /* bounce: reactant separation vector, length, scale */
  double pvalue = -1, rv1[3] = {0, 0, 0}, rv2[3] = {0, 0, 0}; /* bounce by reflection: contact fraction, ends */
  omol* prods[O_MAXPRD];
  if (order == 0) {
    for (d = 0; d < dim; d++) rxnpos[d] = pos0[d];
    rxnbox = sim->nbox ? osim_pos2box(sim, rxnpos) : -1;
  } else if (order == 1) {
    for (d = 0; d < dim; d++) rxnpos[d] = m1->pos[d];
    rxnbox = m1->box;
  } else {
    double dc1 = sim->difc[m1->ident][m1->mstate];
    double dc2 = sim->difc[m2->ident][m2->mstate];
    double x = (dc1 == 0 && dc2 == 0) ? 0.5 : dc2 / (dc1 + dc2);
    for (d = 0; d < dim; d++) rxnpos[d] = x * m1->pos[d] + (1.0 - x) * m2->pos[d];
    rxnbox = (dc1 <= dc2) ? m1->box : m2->box;
  }
  for (prd = 0; prd < rxn->nprod; prd++) {
    omol* mp = osim_getnextmol(sim);
    if (!mp) return 1;
    prods[prd] = mp;
    mp->ident = rxn->prd[prd];
    mp->mstate = rxn->prdstate[prd];
    mp->list = sim->listlookup[mp->ident][mp->mstate];
    mp->box = rxnbox;
    mp->pnl = mp->pnlx = -1;
    if (rxn->rparamt == ORP_BOUNCE && order == 2) {
      /* bounce (smolreact.c:2753-2805): the products keep the reactants'
         serial numbers; by reflection (rparam -2) their paths are reflected
         at the moment of contact, else they move apart along the line joining
         them (rparam >= 0: to distance rparam; otherwise reflected off the
         binding radius) */
      if (prd == 0) {
        pvalue = -1;
        if (rxn->rparam == -2) { /* reflection method: the pair's motion reflected at contact */
          pvalue = omath_sphere_reflect_sphere(m1->posx, m1->pos, m2->posx, m2->pos, dim, rxn->bindrad2, rv1, rv2);
          if (pvalue > 0 && pvalue < 1)
            for (d = 0; d < dim; d++) mp->pos[d] = rv1[d];
          else
            pvalue = -1; /* failed: the other methods */
        }
        if (pvalue == -1) {
          bdist = 0;
          for (d = 0; d < dim; d++) {
            bv[d] = m2->pos[d] - m1->pos[d];
            bdist += bv[d] * bv[d];
          }
          if (bdist == 0) {
            bdist = sqrt(rxn->bindrad2);
            bv[0] = bdist;
          } else
            bdist = sqrt(bdist);
          if (rxn->rparam >= 0) bx = 1.0 / bdist - (1.0 / rxn->rparam);
          else bx = 2.0 / bdist - 2.0 / (rxn->prdpos[0][0] + rxn->prdpos[1][0]);
          for (d = 0; d < dim; d++) mp->pos[d] = m1->pos[d] - bx * rxn->prdpos[0][0] * bv[d];
        }
        mp->box = m1->box;
        mp->serno = m1->serno;
        for (d = 0; d < dim; d++) {
          mp->posx[d] = m1->pos[d];
          mp->posoffset[d] = m1->posoffset[d];
        }
      } else if (prd == 1) {
        mp->box = m2->box;
        mp->serno = m2->serno;
        if (pvalue != -1)
          for (d = 0; d < dim; d++) mp->pos[d] = rv2[d];
        else
          for (d = 0; d < dim; d++) mp->pos[d] = m2->pos[d] + bx * rxn->prdpos[1][0] * bv[d];
        for (d = 0; d < dim; d++) {
          mp->posx[d] = m2->pos[d];
          mp->posoffset[d] = m2->posoffset[d];
        }
      } else {
        for (d = 0; d < dim; d++) mp->posx[d] = mp->pos[d] = rxnpos[d];
      }
      continue;
    }
    if (rxn->rparamt == ORP_CONFSPREAD && order == 2) {
      /* confspread (smolreact.c:2736-2751): the first two products take the
         reactants' places and serial numbers, others the reaction position */
      if (prd < 2) {
        omol* from = prd == 0 ? m1 : m2;
        mp->box = from->box;
        mp->serno = from->serno;
        for (d = 0; d < dim; d++) {
          mp->posx[d] = mp->pos[d] = from->pos[d];
          mp->posoffset[d] = from->posoffset[d];
        }
      } else {
        for (d = 0; d < dim; d++) mp->posx[d] = mp->pos[d] = rxnpos[d];
      }
      continue;
    }
    for (d = 0; d < dim; d++) mp->posx[d] = rxnpos[d];
    for (d = 0; d < dim && rxn->prdpos[prd][d] == 0; d++)
      ;
    if (d != dim) {
      if (rxn->rparamt == ORP_FIXED) {
        for (d = 0; d < dim; d++) v1[d] = rxn->prdpos[prd][d];
      } else if (sim->rngmode == ORNG_COUNTER) {
        /* counter stream: offsets along x turn with one random unit vector;
           when any product's offset leaves the x axis, all offsets of the
           event turn with one random rotation (smolrng_rotation, 3D only) */
        if (!calc) {
          int p2, d2;
          for (p2 = 0; p2 < rxn->nprod; p2++)
            for (d2 = 1; d2 < dim; d2++)
              if (rxn->prdpos[p2][d2] != 0) offaxis = 1;
          if (offaxis && dim != 3) return 12;
          if (offaxis) smolrng_rotation(sim->key0, sim->key1, m1->serno, sim->stepindex, m3);
          else smolrng_unitvec(sim->key0, sim->key1, m1->serno, sim->stepindex, dim, u);
          calc = 1;
        }
        if (offaxis) {
          for (d = 0; d < 3; d++) {
            int j;
            v1[d] = 0;
            for (j = 0; j < 3; j++) v1[d] += m3[3 * d + j] * rxn->prdpos[prd][j];
          }
        } else {
          for (d = 0; d < dim; d++) v1[d] = rxn->prdpos[prd][0] * u[d];
        }
      } else if (dim == 1) {
        if (!calc) {
          m3[0] = (osim_randU32(sim) & 1) ? 1 : -1;
          calc = 1;
        }
        v1[0] = m3[0] * rxn->prdpos[prd][0];
      } else if (dim == 2) {
        if (!calc) {
          omath_dircos2(m3, osim_randCOD(sim) * (2 * O_PI - 0) + 0);
          calc = 1;
        }
        v1[0] = 0;
        v1[0] += m3[0] * rxn->prdpos[prd][0];
        v1[0] += m3[1] * rxn->prdpos[prd][1];
        v1[1] = 0;
        v1[1] += m3[2] * rxn->prdpos[prd][0];
        v1[1] += m3[3] * rxn->prdpos[prd][1];
      } else {
        if (!calc) {
          double th = acos(1.0 - 2.0 * o_randCCD(sim));
          double ph = osim_randCOD(sim) * (2 * O_PI - 0) + 0;
          double ch = osim_randCOD(sim) * (2 * O_PI - 0) + 0;
          omath_dircos(m3, th, ph, ch);
          calc = 1;
        }
        for (d = 0; d < 3; d++) {
          int j;
          v1[d] = 0;
          for (j = 0; j < 3; j++) v1[d] += m3[3 * d + j] * rxn->prdpos[prd][j];
        }
      }
      for (d = 0; d < dim; d++) mp->pos[d] = mp->posx[d] + v1[d];
    } else {
      for (d = 0; d < dim; d++) mp->pos[d] = mp->posx[d];
    }
  }
  o_logevent(sim, order, r, m1, m2, prods, rxn->nprod, rxnpos);
  if (m1) osim_molkill(sim, m1, ll1, mi1);
  if (m2) osim_molkill(sim, m2, ll2, mi2);
  return 0;
}

/* zeroreact (smolreact.c:2905-2993), system-wide reactions: a Poisson number
   of events of mean prob, each at a uniform position in the open system box
   (systemrandpos, smolwall.c:42-46) creating the products there.  Restated
   for the counter stream only: the count is the chunked multiplication
   method of smolrng_poisson_chunk, position draws are keyed by reaction and
   event. */
static int o_zeroreact(osim* sim) {
  orxnss* ss = &sim->rxnss[0];
  int r, d;
  if (!ss->present) return 0;
  if (sim->rngmode != ORNG_COUNTER) return 2;
  for (r = 0; r < ss->totrxn; r++) {
    orxn* rxn = ss->rxn[r];
    int nmol = 0, m, j, k;
    double mu, g;
    if (!(rxn->prob > 0)) continue;
    m = (int)ceil(rxn->prob / SMOLRNG_ZERO_CHUNK);
    mu = rxn->prob / m;
    g = exp(-mu);
    for (j = 0; j < m; j++) nmol += smolrng_poisson_chunk(sim->key0, sim->key1, (uint32_t)r, (uint32_t)j,
                                                          sim->stepindex, g);
    for (k = 0; k < nmol; k++) {
      double pos[O_DIMMAX];
      smolrng4 b = smolrng_philox(sim->key0, sim->key1, (uint32_t)r, (uint32_t)k, sim->stepindex, SMOLRNG_ZEROPOS);
      for (d = 0; d < sim->dim; d++) {
        double lo = sim->wpos[2 * d], hi = sim->wpos[2 * d + 1];
        pos[d] = lo + (hi - lo) * smolrng_u01open(b.v[d]);
      }
      if (o_doreact(sim, 0, r, NULL, NULL, -1, -1, -1, -1, pos)) return 1;
    }
  }
  return 0;
}

/* unireact (smolreact.c:2997-3054) */
static int o_unireact(osim* sim) {
  orxnss* ss = &sim->rxnss[1];
  int ll, m, j;
  if (!ss->present) return 0;
  for (ll = 0; ll < sim->nlist; ll++) {
    int nmol;
    if (!ss->rxnmollist[ll]) continue;
    nmol = sim->nl[ll];
    for (m = 0; m < nmol; m++) {
      omol* mp = sim->live[ll][m];
      int i = mp->ident, ms = mp->mstate;
      smolrng4 rb;
      int blk = -1;
      for (j = 0; j < ss->nrxn[i]; j++) {
        int r = ss->table[i][j];
        orxn* rxn = ss->rxn[r];
        int coin;
        if (!rxn->permit[ms]) continue;
        if (sim->rngmode == ORNG_SFMT) coin = osim_randCOD(sim) < rxn->prob;
        else {
          if (blk != j / 4) {
            blk = j / 4;
            rb = smolrng_mol(sim->key0, sim->key1, mp->serno, sim->stepindex, SMOLRNG_UNIRXN, (uint32_t)blk);
          }
          coin = smolrng_u01(rb.v[j % 4]) < rxn->prob;
        }
        if (!coin) continue;
        if (mp->ident == 0) continue;
        if (o_doreact(sim, 1, r, mp, NULL, ll, m, -1, -1, NULL)) return 1;
        sim->eventcount[0]++;
        j = ss->nrxn[i];
      }
    }
  }
  return 0;
}

/* wallcalcdist2 (smolwall.c:100-116) */
static double o_wallcalcdist2(osim* sim, const double* p1, const double* p2, int wpcode, double* vect) {
  double dist2 = 0;
  int d;
  for (d = 0; d < sim->dim; d++) {
    int c = (wpcode >> (2 * d)) & 3;
    if (c == 0) vect[d] = p2[d] - p1[d];
    else if (c == 1) vect[d] = p2[d] - p1[d] - (sim->wpos[2 * d + 1] - sim->wpos[2 * d]);
    else vect[d] = p2[d] - p1[d] + (sim->wpos[2 * d + 1] - sim->wpos[2 * d]);
    dist2 += vect[d] * vect[d];
  }
  return dist2;
}

int osurf_rxnXsurface_j(osim* sim, omol* m1, omol* m2, int j);

/* morebireact (smolreact.c:3109-3158), solution molecules */
static int o_morebireact(osim* sim, int r, omol* mp1, omol* mp2, int ll1, int m1, int ll2, int et,
                         const double* vect) {
  orxn* rxn = sim->rxnss[2].rxn[r];
  omol *mA, *mB;
  int swap, d, ms;
  if (mp1->ident == rxn->rct[0]) {
    mA = mp1;
    mB = mp2;
    swap = 0;
  } else {
    mA = mp2;
    mB = mp1;
    swap = 1;
  }
  ms = mA->mstate * O_MSMAX1 + mB->mstate;
  if (!rxn->permit[ms]) return 0;
  if (et == 3 && rxn->rparamt != ORP_CONFSPREAD) {
    if (sim->difc[mp1->ident][mp1->mstate] < sim->difc[mp2->ident][mp2->mstate]) {
      for (d = 0; d < sim->dim; d++) {
        double shift = mp1->pos[d] + vect[d] - mp2->pos[d];
        mp2->posoffset[d] -= shift;
        mp2->pos[d] += shift;
        mp2->posx[d] += shift;
      }
    } else {
      for (d = 0; d < sim->dim; d++) {
        double shift = mp2->pos[d] - vect[d] - mp1->pos[d];
        mp1->posoffset[d] -= shift;
        mp1->pos[d] += shift;
        mp1->posx[d] += shift;
      }
    }
  }
  sim->eventcount[et]++;
  if (!swap) return o_doreact(sim, 2, r, mA, mB, ll1, m1, ll2, -1, NULL);
  return o_doreact(sim, 2, r, mA, mB, ll2, -1, ll1, m1, NULL);
}

static int o_bicoin(osim* sim, orxn* rxn, omol* a, omol* b, int j) {
  if (rxn->prob == 1) return 1;
  if (sim->rngmode == ORNG_SFMT) return osim_randCOD(sim) < rxn->prob;
  {
    smolrng4 rr = smolrng_philox(sim->key0, sim->key1, (uint32_t)a->serno, (uint32_t)b->serno, sim->stepindex,
                                 SMOLRNG_BIRXN | ((uint32_t)j << 8));
    return smolrng_u01(rr.v[0]) < rxn->prob;
  }
}

/* bireact (smolreact.c:3201-3296) */
static int o_bireact(osim* sim, int neigh) {
  orxnss* ss = &sim->rxnss[2];
  int ll1, ll2, m1, m2, j, d, b2;
  if (!ss->present) return 0;
  for (ll1 = 0; ll1 < sim->nlist; ll1++)
    for (ll2 = ll1; ll2 < sim->nlist; ll2++) {
      if (!ss->rxnmollist[ll1 * O_MAXLIST + ll2]) continue;
      for (m1 = 0; m1 < sim->nl[ll1]; m1++) {
        omol* mp1 = sim->live[ll1][m1];
        obox* bp = &sim->boxes[mp1->box];
        if (!neigh) {
          omol** ml2 = bp->mol[ll2];
          int nm2 = bp->nmol[ll2];
          for (m2 = 0; m2 < nm2 && ml2[m2] != mp1; m2++) {
            omol* mp2 = ml2[m2];
            int i = mp1->ident * O_MAXSPECIES + mp2->ident;
            for (j = 0; j < ss->nrxn[i]; j++) {
              orxn* rxn = ss->rxn[ss->table[i][j]];
              double dist2 = 0;
              for (d = 0; d < sim->dim; d++) dist2 += (mp1->pos[d] - mp2->pos[d]) * (mp1->pos[d] - mp2->pos[d]);
              if (dist2 <= rxn->bindrad2 && o_bicoin(sim, rxn, mp1, mp2, j) &&
                  (mp1->mstate != OMS_SOLN || mp2->mstate != OMS_SOLN || !osurf_rxnXsurface_j(sim, mp1, mp2, j)) &&
                  mp1->ident != 0 && mp2->ident != 0) {
                if (o_morebireact(sim, ss->table[i][j], mp1, mp2, ll1, m1, ll2, 1, NULL)) return 1;
                if (mp1->ident == 0) {
                  j = ss->nrxn[i];
                  m2 = nm2;
                }
              }
            }
          }
        } else {
          int bmax = (ll1 != ll2) ? bp->nneigh : bp->midneigh;
          for (b2 = 0; b2 < bmax; b2++) {
            obox* bp2 = &sim->boxes[bp->neigh[b2]];
            omol** ml2 = bp2->mol[ll2];
            int nm2 = bp2->nmol[ll2];
            int wpcode = bp->wpneigh ? bp->wpneigh[b2] : 0;
            for (m2 = 0; m2 < nm2; m2++) {
              omol* mp2 = ml2[m2];
              int i = mp1->ident * O_MAXSPECIES + mp2->ident;
              for (j = 0; j < ss->nrxn[i]; j++) {
                orxn* rxn = ss->rxn[ss->table[i][j]];
                double dist2, vect[O_DIMMAX];
                int ok;
                if (wpcode) {
                  dist2 = o_wallcalcdist2(sim, mp1->pos, mp2->pos, wpcode, vect);
                  ok = dist2 <= rxn->bindrad2 && o_bicoin(sim, rxn, mp1, mp2, j) && mp1->ident != 0 &&
                       mp2->ident != 0;
                } else {
                  dist2 = 0;
                  for (d = 0; d < sim->dim; d++)
                    dist2 += (mp1->pos[d] - mp2->pos[d]) * (mp1->pos[d] - mp2->pos[d]);
                  ok = dist2 <= rxn->bindrad2 && o_bicoin(sim, rxn, mp1, mp2, j) &&
                       (mp1->mstate != OMS_SOLN || mp2->mstate != OMS_SOLN || !osurf_rxnXsurface_j(sim, mp1, mp2, j)) &&
                       mp1->ident != 0 && mp2->ident != 0;
                }
                if (ok) {
                  if (o_morebireact(sim, ss->table[i][j], mp1, mp2, ll1, m1, ll2, wpcode ? 3 : 2, vect)) return 1;
                  if (mp1->ident == 0) {
                    j = ss->nrxn[i];
                    m2 = nm2;
                    b2 = bmax;
                  }
                }
              }
            }
          }
        }
      }
    }
  return 0;
}

int osurf_checksurfaces(osim* sim, int ll, int reborn);
int osurf_checksurfacebound(osim* sim, int ll);

/* simulatetimestep (smolsim.cpp:2607-2690) */
int osim_step(osim* sim) {
  int ll, er;
  er = osim_update(sim);
  if (er) return 8;
  sim->nlog = 0;
  if (o_diffuse(sim)) return 9;
  sim->inject = NULL;
  sim->ninject = 0;
  if (sim->srfss) { /* an interaction the oracle does not restate ends the step with 13 */
    for (ll = 0; ll < sim->nlist; ll++)
      if (osurf_checksurfaces(sim, ll, 0) < 0) return 13;
    for (ll = 0; ll < sim->nlist; ll++)
      if (osurf_checksurfacebound(sim, ll) < 0) return 13;
  } else {
    for (ll = 0; ll < sim->nlist; ll++)
      if (sim->diffuselist[ll]) o_checkwalls(sim, ll, 0);
  }
  o_reassign(sim, 1, 0);
  o_molsort(sim, 0);
  er = o_zeroreact(sim);
  if (er) return er == 2 ? 12 : 3;
  if (o_unireact(sim)) return 4;
  if (o_bireact(sim, 0)) return 5;
  if (o_bireact(sim, 1)) return 5;
  o_molsort(sim, 0);
  if (sim->srfss) {
    for (ll = 0; ll < sim->nlist; ll++)
      if (osurf_checksurfaces(sim, ll, 1) < 0) return 13;
  } else {
    for (ll = 0; ll < sim->nlist; ll++) o_checkwalls(sim, ll, 1);
  }
  o_reassign(sim, 0, 1);
  sim->time += sim->dt;
  sim->stepindex++;
  /* simdocommands (smolsim.cpp:2571-2581), called at the end of every step
     (:2679), sorts the lists once more: molecules killed or moved by the
     reborn wall or surface checks leave their lists before the next step's
     diffusion (run-time commands themselves are restated in ocmd.c) */
  o_molsort(sim, 0);
  if (sim->time >= sim->tmax) return 1;
  if (sim->time >= sim->tbreak) return 10;
  return 0;
}

int osim_run(osim* sim, long maxsteps) {
  long k;
  int er = 0;
  for (k = 0; k < maxsteps || maxsteps < 0; k++) {
    er = osim_step(sim);
    if (er) break;
  }
  return er;
}

/* ------------------------------------------------------------------------ */
/* inspection                                                                */

double osim_time(osim* sim) { return sim->time; }
int osim_nlist(osim* sim) { return sim->nlist; }
int osim_list_count(osim* sim, int ll) { return (ll >= 0 && ll < sim->nlist) ? sim->nl[ll] : -1; }

/* live list ll in order: ident, mstate, pos[dim], serno, box */
int osim_get_list(osim* sim, int ll, int* ident, int* mstate, double* pos, uint64_t* serno, int* box) {
  int m, d;
  if (ll < 0 || ll >= sim->nlist) return -1;
  for (m = 0; m < sim->nl[ll]; m++) {
    omol* mp = sim->live[ll][m];
    if (ident) ident[m] = mp->ident;
    if (mstate) mstate[m] = mp->mstate;
    if (pos)
      for (d = 0; d < sim->dim; d++) pos[m * sim->dim + d] = mp->pos[d];
    if (serno) serno[m] = mp->serno;
    if (box) box[m] = mp->box;
  }
  return sim->nl[ll];
}

int osim_count_species(osim* sim, int ident) {
  int ll, m, c = 0;
  for (ll = 0; ll < sim->nlist; ll++)
    for (m = 0; m < sim->nl[ll]; m++)
      if (sim->live[ll][m]->ident == ident) c++;
  for (m = sim->topd; m < sim->nd; m++)
    if (sim->dead[m]->ident == ident) c++;
  return c;
}

static const char* o_ms2string(int ms) {
  static const char* s[] = {"solution", "front", "back", "up", "down", "bsoln", "all", "none", "some"};
  return (ms >= 0 && ms <= 8) ? s[ms] : "none";
}

/* cmdlistmols (smolcmd.c:2044-2075) via molscancmd (smolmolec.c:1178-1205):
   "name(state) x y z serno" with %g, live lists then resurrected molecules. */
long osim_listmols(osim* sim, char* buf, long buflen) {
  long n = 0;
  int ll, m, d;
  char line[512];
  for (ll = -1; ll < sim->nlist; ll++) {
    int lo, hi;
    omol** ml;
    if (ll < 0) continue;
    lo = 0;
    hi = sim->nl[ll];
    ml = sim->live[ll];
    for (m = lo; m < hi; m++) {
      omol* mp = ml[m];
      int k;
      if (mp->ident <= 0) continue;
      k = snprintf(line, sizeof(line), "%s(%s)", sim->spname[mp->ident], o_ms2string(mp->mstate));
      for (d = 0; d < sim->dim; d++) k += snprintf(line + k, sizeof(line) - k, " %g", mp->pos[d]);
      k += snprintf(line + k, sizeof(line) - k, " %llu\n", (unsigned long long)mp->serno);
      if (buf && n + k < buflen) memcpy(buf + n, line, k);
      n += k;
    }
  }
  for (m = sim->topd; m < sim->nd; m++) {
    omol* mp = sim->dead[m];
    int k;
    if (mp->ident <= 0) continue;
    k = snprintf(line, sizeof(line), "%s(%s)", sim->spname[mp->ident], o_ms2string(mp->mstate));
    for (d = 0; d < sim->dim; d++) k += snprintf(line + k, sizeof(line) - k, " %g", mp->pos[d]);
    k += snprintf(line + k, sizeof(line) - k, " %llu\n", (unsigned long long)mp->serno);
    if (buf && n + k < buflen) memcpy(buf + n, line, k);
    n += k;
  }
  if (buf && n < buflen) buf[n] = '\0';
  return n;
}

int osim_rxn_info(osim* sim, int order, int r, double* out) {
  orxn* rxn;
  if (order < 1 || order > 2 || r < 0 || r >= sim->rxnss[order].totrxn) return 1;
  rxn = sim->rxnss[order].rxn[r];
  out[0] = rxn->bindrad2;
  out[1] = rxn->prob;
  out[2] = rxn->unbindrad;
  out[3] = rxn->prdpos[0][0];
  out[4] = rxn->nprod > 1 ? rxn->prdpos[1][0] : 0;
  out[5] = rxn->rparamt;
  out[6] = rxn->rparam;
  return 0;
}

int osim_box_info(osim* sim, int* side, double* bmin, double* bsize) {
  int d;
  for (d = 0; d < sim->dim; d++) {
    side[d] = sim->side[d];
    bmin[d] = sim->bmin[d];
    bsize[d] = sim->bsize[d];
  }
  return sim->nbox;
}

int osim_box_neighbors(osim* sim, int b, int* neigh, int* wp, int* midneigh) {
  obox* bp;
  int k;
  if (b < 0 || b >= sim->nbox) return -1;
  bp = &sim->boxes[b];
  for (k = 0; k < bp->nneigh; k++) {
    neigh[k] = bp->neigh[k];
    wp[k] = bp->wpneigh ? bp->wpneigh[k] : 0;
  }
  *midneigh = bp->midneigh;
  return bp->nneigh;
}

/* reaction events of the last step: rows of 8 int64:
   order, r, reactant sernos (2), product sernos (2), 0, 0 */
int osim_rxnlog(osim* sim, int64_t* rows, int maxrows) {
  int k;
  for (k = 0; k < sim->nlog && k < maxrows; k++) {
    orxnevent* e = &sim->log[k];
    rows[8 * k + 0] = e->order;
    rows[8 * k + 1] = e->r;
    rows[8 * k + 2] = (int64_t)e->rct[0];
    rows[8 * k + 3] = (int64_t)e->rct[1];
    rows[8 * k + 4] = (int64_t)e->prd[0];
    rows[8 * k + 5] = (int64_t)e->prd[1];
    rows[8 * k + 6] = 0;
    rows[8 * k + 7] = 0;
  }
  return sim->nlog;
}

/* reaction positions of the logged events (dim values each) */
int osim_rxnlog_pos(osim* sim, double* pos, int maxrows) {
  int k, d;
  for (k = 0; k < sim->nlog && k < maxrows; k++)
    for (d = 0; d < sim->dim; d++) pos[sim->dim * k + d] = sim->log[k].pos[d];
  return sim->nlog;
}

uint64_t osim_next_serno(osim* sim) { return sim->serno; }

const double* osim_gausstable(osim* sim, int flat) { return flat ? sim->gtable_flat : sim->gtable; }
int osim_ngauss(osim* sim) { return sim->ngauss; }
