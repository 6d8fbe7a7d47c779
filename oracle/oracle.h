/* CPU oracle for the MI355X Smoldyn engine -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's per-timestep particle path
 * (ssandrews/Smoldyn source/Smoldyn/smol*.c, source/libSteve; each
 * function cites the file:line it follows).  Only tests/, the smoke check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may load it, and only as
 * the checker or the timed CPU baseline.  The product (libsmoldyn_amd.so)
 * never links, loads or calls it.
 *
 * Parity status of each part (and how it is pinned) is listed in DESIGN.md;
 * every function here is pinned by at least one of: the reference's own
 * regression outputs (examples/S95_regression/output_275), the compiled
 * reference libSteve sources (oracle/_ref), closed forms, or brute force.
 */
#ifndef SMOLDYN_ORACLE_H
#define SMOLDYN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osim osim;

osim* osim_new(int dim, const double* low, const double* high);
void osim_free(osim* sim);
int osim_set_wall(osim* sim, int d, int highside, char type);
void osim_set_times(osim* sim, double tstart, double tstop, double dt);
void osim_set_seed(osim* sim, long seed);
void osim_set_rngmode(osim* sim, int mode, uint64_t key);
int osim_set_partitions(osim* sim, const char* method, double value);
void osim_set_accuracy(osim* sim, double accur);
int osim_add_species(osim* sim, const char* name);
int osim_species_index(osim* sim, const char* name);
void osim_set_difc(osim* sim, int i, int ms, double difc);
int osim_add_list(osim* sim, const char* name);
int osim_set_list(osim* sim, int i, int ms, int ll);
int osim_add_mols(osim* sim, int n, int ident, const double* lo, const double* hi);
int osim_add_mols_at(osim* sim, int n, int ident, const double* pos);
int osim_add_reaction(osim* sim, const char* name, int order, const int* rct, const int* rctstate, int nprod,
                      const int* prd, const int* prdstate, double rate);
int osim_set_revparam(osim* sim, int order, int r, int rparamt, double rparam);
int osim_set_rxn_prob(osim* sim, int order, int r, double prob);
int osim_update(osim* sim);
int osim_inject(osim* sim, const double* g, long n);
int osim_step(osim* sim);
int osim_run(osim* sim, long maxsteps);

double osim_time(osim* sim);
int osim_nlist(osim* sim);
int osim_list_count(osim* sim, int ll);
int osim_get_list(osim* sim, int ll, int* ident, int* mstate, double* pos, uint64_t* serno, int* box);
int osim_count_species(osim* sim, int ident);
long osim_listmols(osim* sim, char* buf, long buflen);
int osim_rxn_info(osim* sim, int order, int r, double* out);
int osim_box_info(osim* sim, int* side, double* bmin, double* bsize);
int osim_box_neighbors(osim* sim, int b, int* neigh, int* wp, int* midneigh);
int osim_rxnlog(osim* sim, int64_t* rows, int maxrows);
uint64_t osim_next_serno(osim* sim);
const double* osim_gausstable(osim* sim, int flat);
int osim_ngauss(osim* sim);

/* surfaces (osurf.c).  Shapes: 0 rect, 1 tri, 2 sph, 5 disk; faces 0 front,
   1 back, 2 none, 3 both; actions as the reference's SrfAction. */
int osim_add_surface(osim* sim, const char* name);
int osim_set_surface_action(osim* sim, int s, int ident, int ms, int face, int act, int newident);
int osim_set_surface_rate(osim* sim, int s, int ident, int ms, int ms1, int ms2, int newident, double value,
                          int which);
int osim_add_panel(osim* sim, int s, int ps, const char* rectaxis, const double* params);
int osim_set_panel_jump(osim* sim, int p1, int f1, int p2, int f2, int bidirectional);
int osim_box_panels(osim* sim, int b, int* out, int maxn);
/* geometry restatements, exported for the tests against the reference */
double osurf_geo_spherenormal(const double* cent, const double* pt, int front, int dim, double* ans);
int osurf_geo_ptinslab(const double* pt1, const double* pt2, const double* test, int dim);
double osurf_geo_linexsphs(const double* pt1, const double* pt2, const double* cent, double rad, int dim,
                           double* crss2ptr, double* nrdistptr, double* nrposptr);
int osurf_geo_rectxaabb3(const double* r1, const double* r2, const double* r3, const double* r4,
                         const double* bpt1, const double* bpt2);
int osurf_geo_sphsxaabb3(const double* cent, double rad, const double* bpt1, const double* bpt2);
int osurf_geo_diskxaabb3(const double* cent, double rad, const double* norm, const double* bpt1,
                         const double* bpt2);
double osurf_desorbdist_y(double step, double y, int revads);
/* osurfparam.c: rates -> collision probabilities (SurfaceParam.c) */
double osurf_surfaceprob(double k1, double k2, double dt, double difc, double* p2ptr, int algo);
int osurf_transmit(double kap1, double kap2, double difc1, double difc2, double dt, double* p1ptr, double* p2ptr);
double osurf_lookupirrevadsorb(double value);
double osurf_lookuprevads(double value1, double value2, double* ans2ptr);
double osurf_experfcm1(double x);
int osurf_geo_linexaabb2(const double* pt1, const double* pt2, const double* norm, const double* bpt1,
                         const double* bpt2);
int osurf_geo_circlexaabb2(const double* cent, double rad, const double* bpt1, const double* bpt2);
double osurf_geo_linenormal(const double* pt1, const double* pt2, double* ans);
double osurf_geo_trinormal(const double* pt1, const double* pt2, const double* pt3, double* ans);
double osurf_geo_unitcross(const double* v1start, const double* v1end, const double* v2start, const double* v2end,
                           double* ans);
int osurf_geo_ptintriangle2(const double (*point)[3], const double* test);
double osurf_geo_linenormal3d(const double* pt1, const double* pt2, const double* point, double* ans);
double osurf_geo_linexcyls(const double* pt1, const double* pt2, const double* cp1, const double* cp2, double rad,
                           double* crss2ptr, double* nrdistptr, double* nrposptr);
int osurf_geo_linexaabb(const double* pt1, const double* pt2, const double* bpt1, const double* bpt2, int dim,
                        int infline);
int osurf_geo_cylisxaabb3(const double* pt1, const double* pt2, double rad, const double* bpt1, const double* bpt2);
int osurf_geo_cylsxaabb3(const double* pt1, const double* pt2, double rad, const double* bpt1, const double* bpt2);
int osurf_geo_hemisxaabb3(const double* cent, double rad, const double* outvect, const double* bpt1,
                          const double* bpt2);
int osurf_geo_trixaabb3(const double* pt1, const double* pt2, const double* pt3, const double* norm,
                        const double* bpt1, const double* bpt2);

/* rxnparam restatement, exported for the rate-table tests */
double orxn_numrxnrate(double step, double a, double b);
double orxn_numrxnrateprob(double step, double a, double b, double prob);
double orxn_actrxnrate(double step, double a, double prob);
double orxn_bindingradius(double rate, double dt, double difc, double b, int rel);
double orxn_bindingradiusprob(double rate, double dt, double difc, double b, int rel, double chi, double* probptr);
double orxn_unbindingradius(double pgem, double dt, double difc, double a);
void omath_randtable(double* a, int n);
double omath_erfn(double x);

int osim_set_confspread(osim* sim, int r, double radius);
int osim_set_bindrad(osim* sim, int r, double radius);
int osim_rxnlog_pos(osim* sim, double* pos, int maxrows);
int osim_set_prdpos(osim* sim, int order, int r, int prd, const double* pos);
void osim_rng_rotation(uint32_t k0, uint32_t k1, uint64_t serial, uint32_t step, double* m);

/* run-time command scheduler restatement (ocmd.c) driven by a script */
int ocmd_timeline(const char* script, int maxev, double* et, int* ec, long long* ei);

#ifdef __cplusplus
}
#endif
#endif
