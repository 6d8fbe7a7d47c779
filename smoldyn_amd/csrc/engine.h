// Device engine of the MI355X Smoldyn particle step: data layout shared by the
// host library (capi.cpp) and the HIP kernels (engine.hip).
//
// Molecules live in structure-of-arrays form in HBM, one contiguous segment
// per molecule list, in the reference's live-list order (the order that
// source/Smoldyn/smolmolec.c:2943-3023 molsort maintains).  Keeping that
// order on the device is what lets the greedy pair resolution of bireact
// (smolreact.c:3201-3296) be reproduced exactly.
#pragma once

#include <stdint.h>

// Device arrays reached through pointers kept in device memory (a
// simulation's view, its surface tables) are declared in the global address
// space for the device compilation: loads and stores through them are global
// memory instructions, whose waits are separate from LDS waits (through a
// generic pointer they would be flat instructions, which every LDS wait also
// waits for).  Host code sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define SMOLAMD_GLOBAL __attribute__((address_space(1)))
#else
#define SMOLAMD_GLOBAL
#endif

namespace smolamd {

constexpr int kMaxSpecies = 32;
constexpr int kMaxList = 8;
constexpr int kMaxRxnPerSet = 4;  // reactions per reactant (pair)
constexpr int kMaxRxn = 32;
constexpr int kMaxProd = 4;
constexpr uint64_t kNoKey = ~0ull;

// wall kinds, per wall 2*d+side
enum WallKind : int { kWallTransparent = 0, kWallReflect = 1, kWallPeriodic = 2, kWallAbsorb = 3 };

struct DevRxn {
  int order;
  int rct0;        // species of the first reactant (rctident[0])
  int rct1;        // species of the second reactant (order 2)
  int nprod;
  int prd[kMaxProd];
  int prdlist[kMaxProd];   // molecule list of each product
  double bindrad2;
  double prob;
  double prdpos[kMaxProd];  // axial product offset (prdpos[prd][0])
  double prdvec[kMaxProd][3];  // product offset vectors of fixed placement
  int rotate;               // 1 if any product offset is nonzero
  int permit_soln;          // permitted for solution reactant(s)
  int confspread;           // products 0 and 1 take the reactants' places and serial numbers
  int fixed;                // products at the reaction position plus prdvec, not rotated
  int offrot;               // offsets off the x axis: prdvec turned by one random rotation per event
  int bounce;               // products 0 and 1 keep the reactants' serials and move apart (scaling methods)
  double bparam;            // bounce parameter: >= 0 final separation, < 0 reflection off the binding radius
};

constexpr int kMaxRxn0 = 8;  // zeroth-order reactions

struct DevTables {
  int nspecies;
  int nlist;
  int dim;
  int ngauss;
  // zeroth-order reactions: events per step are Poisson of mean prob, drawn in
  // zm chunks of mean prob/zm against zg = exp(-prob/zm); each event makes
  // its product at a uniform position in the open system box [zlo, zhi]
  int nrxn0;
  DevRxn rxn0[kMaxRxn0];
  double zg[kMaxRxn0];
  int zm[kMaxRxn0];
  double zlo[3], zhi[3];
  double difc[kMaxSpecies];
  double difstep[kMaxSpecies];
  int listof[kMaxSpecies];  // list of the solution state
  int nuni[kMaxSpecies];
  int uni[kMaxSpecies][kMaxRxnPerSet];
  int nbi[kMaxSpecies * kMaxSpecies];
  int bi[kMaxSpecies * kMaxSpecies][kMaxRxnPerSet];
  int rxnmol2[kMaxList * kMaxList];
  int pairindex[kMaxList * kMaxList];  // rank of (ll1,ll2), ll1<=ll2, in the reference's loop order
  int diffuselist[kMaxList];
  uint32_t reactmask[kMaxSpecies];  // bit i2 set when species (i, i2) has a bimolecular reaction
  double maxbind2[kMaxSpecies];     // largest squared binding radius over those reactions
  DevRxn rxn1[kMaxRxn];
  DevRxn rxn2[kMaxRxn];
};

struct Geometry {
  double low[3], high[3], width[3];
  int wall[6];
  int side[3];
  int wrapflag[3];  // neighbor wrap flags as neighborZV type 7 reads them
  double bmin[3], bsize[3];
  int nbox;
  // FP32 pre-filter of the pair search: a pair is tested exactly (FP64) when
  // its FP32 distance is within binding radius + prefilter_margin; the margin
  // bounds the FP32 rounding of coordinates up to twice the system extent
  // (molecules beyond that bypass the pre-filter).
  double prefilter_margin;
};

// Device control block: counters updated by the kernels themselves, so a
// whole time step runs without host synchronisation.
struct Control {
  int nl[kMaxList];
  int pfx[kMaxList];       // diffuse-order offset of each list (injection index)
  int nfresh[kMaxList];    // molecules appended since the last step at each list's end: they skip this step's diffusion
  uint64_t serno_next;
  uint32_t step;           // step index, a Philox counter word
  int nedges;
  int nuni;
  int naccept;
  int nlog;
  int err;                 // bit flags, see ErrBits
  int count[kMaxSpecies];  // molecules per species (all states)
  int countf[kMaxSpecies];  // of which bound in the front state
  int countk[kMaxSpecies];  // of which bound in the back state
  int nmove;               // list changes queued by k_diffuse for the first molsort (surfaces)
  int nwall;               // absorbing walls: survivors the second wall pass kills (k_diffuse -> k_resolve)
  int maxedges_seen;
  int rounds;              // matching rounds of the last step
  unsigned long long molsteps;  // sum over steps of the live molecules at step start
  int nbinned;             // records binned in the last step (all simulations of a batch)
};

enum ErrBits : int {
  kErrEdgeOverflow = 1,
  kErrEventOverflow = 2,
  kErrCapacity = 4,
  kErrSortOverflow = 8,
  kErrSurface = 16,  // a surface action the device path does not run (absorption, adsorption of a product, ...)
  kErrHalted = 32,   // stepping stopped after an error (later steps did nothing)
};

// ---------------------------------------------------------------------------
// surfaces (solution-phase molecules; 3D panels of every shape)

constexpr int kMaxSurf = 16;
enum DevPanelShape : int { kPsRect = 0, kPsTri = 1, kPsSph = 2, kPsCyl = 3, kPsHemi = 4, kPsDisk = 5 };
constexpr int kPanelPoints = 6;
enum DevSrfAction : int {
  kSaReflect = 0, kSaTrans = 1, kSaAbsorb = 2, kSaJump = 3, kSaMult = 5, kSaNo = 6, kSaAdsorb = 8, kSaRevdes = 9,
  kSaIrrevdes = 10, kSaFlip = 11
};

struct DevPanel {
  int ps;           // DevPanelShape
  int srf;          // surface index
  int jumpp[2];     // jump destination panel per face (-1 none)
  int jumpf[2];     // destination face
  // rect: corners 0..2; tri: vertices 0..2, outward unit edge normals 3..5;
  // sph/hemi/disk: center, then radius in point[1][0]; hemi: unit vector out
  // of the opening in point[2]; cyl: axis ends 0, 1, radius in point[2][0]
  double point[kPanelPoints][3];
  // rect: (+-1, normal axis, edge axis); tri/disk: unit normal; sph/hemi:
  // (+-1); cyl: front[2] = +-1
  double front[3];
};

// Surface actions of solution-phase molecules, per surface, species and face
// (0 front, 1 back).  A rate-based action (kSaMult) draws one uniform u and
// takes the first state ms with u < cumprob[ms]; outcome[ms] is the action
// that state means on that face and newspec[ms] the species after it.
struct DevSrfAct {
  int act;
  int newspec_fixed;   // species after a fixed action
  double cumprob[6];
  int outcome[6];
  int newspec[6];
};

struct DevSurf {
  int on;                   // surfaces defined: walls are not applied, collisions are checked
  int npnl;
  int listdiffuse[kMaxList];  // lists whose molecules are checked after diffusion
  double epsilon;
  SMOLAMD_GLOBAL const DevPanel* pnl;
  SMOLAMD_GLOBAL const int* bpstart;       // panels of box b: bplist[bpstart[b] .. bpstart[b+1])
  SMOLAMD_GLOBAL const int* bplist;
  SMOLAMD_GLOBAL const DevSrfAct* act;     // [surface][species][face]
  SMOLAMD_GLOBAL const DevSrfAct* actb;    // surface-bound molecules: [surface][species][front/back state], no face
  int listbound[kMaxList];  // lists holding surface-bound states (checksurfacebound)
  int blist[kMaxSpecies][2];  // list of each species in the front / back state (-1: none)
  // no surfaces but absorbing walls: checkwalls with kills instead of the
  // surface checks (the surface variants carry the molsorts a kill needs)
  int walls;
  SMOLAMD_GLOBAL int* wq;  // walls: destination records of survivors the step's second wall pass kills
};

// A list change queued by k_diffuse and applied by the first molsort of the
// step (smolmolec.c:2943-3023): an adsorbed molecule moving to the list of
// its bound state, or a desorbing bound molecule whose list slot dies while
// the record is appended, as a new solution molecule with the same serial,
// to the end of its solution list.
// kMoveKill: a molecule absorbed by a surface; its slot dies in the first molsort
enum MoveKind : int { kMoveList = 1, kMoveDesorb = 2, kMoveKill = 3 };
struct MoveEv {
  int rec;   // record in the source buffer of the step
  int dst;   // record in the destination buffer (k_scatter), -1 when not binned (a first-order reactant)
  int slot;  // list base + live position before the change
  int list;  // list after the change
  int kind;
  int obox;  // box of the molecule in its old list's box lists
  int nbox;  // box it is added to in the new list's box lists
  int ov;    // working order value (k_relist)
};

struct Edge {
  uint64_t key;  // pass | list pair | live pos 1 | neighbor | live pos 2: the reference's loop order
  int s1, s2;    // records (destination buffer) of the iterating molecule and its partner
  int r;         // reaction index (order 2)
  int wrap;      // periodic wrap code of the neighbor box (wallcalcdist2)
};

struct UniEvent {
  int rec;   // record index in the source buffer of the step
  int slot;  // list base + live position
  int r;
  int box;   // virtual box the record's position is stored relative to (FP32-offset records)
};

// One reaction event in creation order (the reaction log): reactant and
// product serial numbers (0 where there is none) and the reaction position
struct LogRow {
  int order, r;
  uint64_t rct[2];
  uint64_t prd[kMaxProd];
  double pos[3];
};

struct DeviceState;  // engine.hip

struct EngineConfig {
  int nlist;
  int cap[kMaxList];
  int maxedges;
  int maxevents;
  double expected_per_box;  // molecules per virtual box per list, for the pair-search tile
  // record precision: 0 = FP64 positions; 1 = FP32 offsets from the FP64
  // origin of the record's virtual box (all arithmetic stays FP64)
  int f32;
  // largest z sub-bin count per virtual box for the pair search (0: automatic)
  int zsub_max;
  // a bimolecular reaction places its products by bounce reflection: keep
  // each record's position before the step's diffusion (one simulation
  // without surfaces)
  int reflect;
};

// Host description of the surfaces handed to engine_create (arrays copied).
struct HostSurf {
  int on;
  int npnl;
  const DevPanel* pnl;
  int nbox;
  const int* bpstart;  // nbox + 1 entries
  const int* bplist;
  int nsrf;
  const DevSrfAct* act;   // nsrf * kMaxSpecies * 2
  const DevSrfAct* actb;  // nsrf * kMaxSpecies * 2
  int listdiffuse[kMaxList];
  int listbound[kMaxList];
  int blist[kMaxSpecies][2];
  double epsilon;
  int walls;  // no panels: absorbing walls run on the surface variants
};

DeviceState* engine_create(const EngineConfig& cfg, const DevTables& tab, const Geometry& geo,
                           const double* gtable, int ngauss, uint32_t key0, uint32_t key1, const HostSurf* surf,
                           char* err, int errlen);
// A batch of n independent simulations stepped together (ensembles): one
// set of launches per step covers all of them, k_resolve runs one workgroup
// per simulation.  Surfaces are supported for n == 1 only.
DeviceState* engine_create_batch(int n, const EngineConfig* cfg, const DevTables* tab, const Geometry* geo,
                                 const uint32_t* key0, const uint32_t* key1, const double* gtable, int ngauss,
                                 const HostSurf* surf, char* err, int errlen);
int engine_num_instances(DeviceState* s);
int engine_upload_list_inst(DeviceState* s, int inst, int ll, int n, const int* ident, const double* pos,
                            const uint64_t* serno);
int engine_set_counters_inst(DeviceState* s, int inst, uint64_t serno_next, uint32_t step, const int* counts);
int engine_get_control(DeviceState* s, int inst, Control* out);
// mstate (optional): MolecState of each molecule (0 solution, 1 front, 2 back)
int engine_get_list_inst(DeviceState* s, int inst, int ll, int* ident, double* pos, uint64_t* serno, int* box,
                         int max, int* mstate = nullptr);
void engine_destroy(DeviceState* s);
// Upload molecules of list ll in live order (host arrays), replacing its content.
int engine_upload_list(DeviceState* s, int ll, int n, const int* ident, const double* pos, const uint64_t* serno);
// Molecules added between steps (one simulation, no surface variant): n
// records appended to the end of list ll in the given order; they join the
// next step's pair search but skip its diffusion.  -3 when the list or the
// record buffer is full.
int engine_append_fresh(DeviceState* s, int ll, int n, const int* ident, const double* pos, const uint64_t* serno);
int engine_set_counters(DeviceState* s, uint64_t serno_next, uint32_t step, const int* counts);
// One reference time step (simulatetimestep).  inject: optional host array of
// 3 Gaussian deviates per diffusing molecule in diffuse order (NULL = Philox).
int engine_step(DeviceState* s, const double* inject, long ninject);
// The same injected step with the deviates staged in a fixed device array
// instead of an allocated buffer: allocates no device memory.  Returns the
// error bits, -2 when ninject is short, -4 when ninject exceeds
// engine_noalloc_inject_max() molecules.
int engine_step_noalloc(DeviceState* s, const double* inject, long ninject);
long engine_noalloc_inject_max();
// Many steps; returns 0 or an error bit set.
int engine_run(DeviceState* s, long nsteps, int use_graph);
// Steps whose kernels carry their own start / stop HIP events (stamped at
// each dispatch's start and end); out[0..9] = average ms per step of
// diffuse, box scan, scatter, first molsort (surfaces), pair screen, exact
// test, resolve, tile binning pass 1, tile offsets, tile binning pass 2.
int engine_profile(DeviceState* s, long nsteps, double* out);
int engine_sync_control(DeviceState* s, Control* out);
// Asynchronous stepping for the command loop: enqueue steps, enqueue a copy of
// the species counters into observation row `row` (kObsStride ints: step,
// count, countf, countk), then wait and read the error bits and the rows.
constexpr int kObsRows = 4096;
constexpr int kObsStride = 1 + 3 * kMaxSpecies;
int engine_enqueue(DeviceState* s, long nsteps);
int engine_observe(DeviceState* s, int row);
int engine_finish(DeviceState* s);
int engine_read_obs(DeviceState* s, int nrow, int* out);
int engine_get_list(DeviceState* s, int ll, int* ident, double* pos, uint64_t* serno, int* box, int max);
int engine_get_log(DeviceState* s, LogRow* rows, int maxrows);
double engine_last_ms(DeviceState* s);

}  // namespace smolamd
