/* CPU oracle simulation state (test infrastructure only; never linked into the
 * product).  Mirrors the reference's data model, source/Smoldyn/smoldyn.h:
 * molecules (:280-293), molecule superstructure with live/dead lists
 * (:295-330), walls, reactions (:389-437), boxes (:625-655). */
#ifndef ORACLE_OSIM_H
#define ORACLE_OSIM_H

#include <stdint.h>

#include "oracle_internal.h"

#define O_DIMMAX 3
#define O_MSMAX 5
#define O_MSMAX1 6
#define O_MAXSPECIES 32
#define O_MAXLIST 8
#define O_MAXRXN 32
#define O_MAXPRD 4
#define O_NAMELEN 64

enum { OMS_SOLN, OMS_FRONT, OMS_BACK, OMS_UP, OMS_DOWN, OMS_BSOLN, OMS_ALL, OMS_NONE, OMS_SOME };
enum {
  ORP_NONE, ORP_IRREV, ORP_CONFSPREAD, ORP_BOUNCE, ORP_PGEM, ORP_PGEMMAX, ORP_PGEMMAXW, ORP_RATIO,
  ORP_UNBINDRAD, ORP_PGEM2, ORP_PGEMMAX2, ORP_RATIO2, ORP_OFFSET, ORP_FIXED
};
enum { ORNG_SFMT = 0, ORNG_COUNTER = 1 };

typedef struct omol {
  uint64_t serno;
  int list;
  double pos[O_DIMMAX], posx[O_DIMMAX], via[O_DIMMAX], posoffset[O_DIMMAX];
  int ident;
  int mstate;
  int box; /* box index, -1 for none */
  int pnl; /* surface panel id, -1 for none */
  int pnlx;
} omol;

typedef struct obox {
  int indx[O_DIMMAX];
  int nneigh, midneigh;
  int* neigh;
  int* wpneigh; /* NULL when no neighbor wraps */
  int nmol[O_MAXLIST], maxmol[O_MAXLIST];
  omol** mol[O_MAXLIST];
} obox;

typedef struct orxn {
  char name[O_NAMELEN];
  int order;
  int rct[2], rctstate[2];
  int permit[O_MSMAX1 * O_MSMAX1];
  int nprod;
  int prd[O_MAXPRD], prdstate[O_MAXPRD];
  double rate;
  int multiplicity;
  double bindrad2, prob, chi;
  int rparamt;
  double rparam, unbindrad;
  double prdpos[O_MAXPRD][O_DIMMAX];
  int disable;
} orxn;

typedef struct orxnss {
  int present;
  int nrxn[O_MAXSPECIES * O_MAXSPECIES];
  int table[O_MAXSPECIES * O_MAXSPECIES][O_MAXRXN];
  int totrxn;
  orxn* rxn[O_MAXRXN];
  int rxnmollist[O_MAXLIST * O_MAXLIST];
} orxnss;

/* One logged reaction event, in the order doreact ran. */
typedef struct orxnevent {
  int order, r;
  uint64_t rct[2];
  uint64_t prd[2];
  double pos[O_DIMMAX];
} orxnevent;

struct osurfss; /* surfaces, osurf.c */

typedef struct osim {
  int dim;
  double wpos[2 * O_DIMMAX];
  char wtype[2 * O_DIMMAX];
  double tmin, tmax, dt, time, tbreak;
  uint32_t stepindex;
  double accur;
  int ok; /* 1 after a successful update with no pending changes */

  /* random streams */
  int rngmode;
  osfmt sfmt;
  uint32_t key0, key1;
  const double* inject; /* injected Gaussian displacements for the next diffuse */
  long ninject;

  /* species */
  int nspecies; /* including the empty species 0 */
  char spname[O_MAXSPECIES][O_NAMELEN];
  double difc[O_MAXSPECIES][O_MSMAX];
  double difstep[O_MAXSPECIES][O_MSMAX];
  int listlookup[O_MAXSPECIES][O_MSMAX];

  /* molecule lists */
  int nlist;
  char listname[O_MAXLIST][O_NAMELEN];
  int listsystem[O_MAXLIST];
  omol** live[O_MAXLIST];
  int maxl[O_MAXLIST], nl[O_MAXLIST], topl[O_MAXLIST], sortl[O_MAXLIST];
  int diffuselist[O_MAXLIST];
  omol** dead;
  int maxd, nd, topd;
  uint64_t serno;
  int ngauss;
  double* gtable;      /* shuffled table (reference stream) */
  double* gtable_flat; /* unshuffled table (counter stream) */

  /* virtual boxes */
  double mpbox, boxsize;
  int side[O_DIMMAX];
  double bmin[O_DIMMAX], bsize[O_DIMMAX];
  int nbox;
  obox* boxes;

  /* reactions, by order */
  orxnss rxnss[3];

  /* surfaces */
  struct osurfss* srfss;

  /* reaction log of the current step */
  orxnevent* log;
  int nlog, maxlog;
  long eventcount[4]; /* rxn1, rxn2intra, rxn2inter, rxn2wrap */
} osim;

/* internal helpers shared with osurf.c */
int osim_pos2box(const osim* sim, const double* pos);
omol* osim_getnextmol(osim* sim);
void osim_molkill(osim* sim, omol* m, int ll, int mi);
void osim_molchangeident(osim* sim, omol* m, int ll, int mi, int i, int ms, int pnl);
double osim_randCOD(osim* sim);
double osim_randCCD(osim* sim);
uint32_t osim_randU32(osim* sim);

#endif
