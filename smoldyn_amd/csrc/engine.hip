// HIP kernels of the MI355X Smoldyn particle step (gfx950).
//
// Data layout.  Molecule records (FP64 position, species, serial number and
// the molecule's position in the reference's live list) are stored in
// virtual-box order, double-buffered, so every per-molecule pass is a
// coalesced HBM stream.  The reference's ordering semantics only depend on
// live-list positions (molsort, source/Smoldyn/smolmolec.c:2943-3023, keeps
// them; reassignmolecs, smolboxes.c:896-954, fills box lists in live order),
// so the live position is kept as a record attribute plus an inverse map
// (list slot -> record) instead of as the storage order.
//
// One reference time step (smolsim.cpp:2607-2690 simulatetimestep) runs as
// six launches and no host synchronisation:
//
//   k_diffuse   diffuse (smolmolec.c:3115-3168), checkwalls (smolwall.c:368-424)
//               or checksurfaces, pos2box (smolboxes.c:67-79), unireact coins
//               (smolreact.c:2997-3054), histogram of (list, box, z sub-bin);
//               one thread per record of the source buffer
//   k_hsum, k_hscan  exclusive scan of the histogram (box lists of
//               reassignmolecs, each box's records ordered by z sub-bin);
//               k_scan2 scans the tile sums in between for large scans
//   k_scatter   counting-sort scatter into the destination buffer + inverse map
//               + packed FP32 payload for the pair search
//   k_pairs     bireact candidates (smolreact.c:3201-3296): same box and half
//               or full neighbor stencil, FP32 screen then the reference's FP64
//               distance test; edges keyed by the reference's loop order
//   k_resolve   one workgroup: greedy matching in key order, products
//               (doreact, smolreact.c:2662-2901), molsort compaction and the
//               dead-list product append order
//
// Arithmetic follows the reference operation by operation and the library is
// built with -ffp-contract=off, so FP64 positions match the CPU path bit for
// bit.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "../../include/smoldyn_amd_rng.h"
#include "engine.h"

namespace smolamd {

#define HIPOK(call)                                                                  \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      fprintf(stderr, "smoldyn_amd: %s failed: %s\n", #call, hipGetErrorString(e_)); \
      return -1;                                                                     \
    }                                                                                \
  } while (0)

constexpr int kBlock = 256;
constexpr int kExactBlocks = 512;  // grid of k_exact (grid-stride over the candidates)
constexpr int kRankBits = 27;  // arrival rank within a histogram bucket; the z sub-bin sits above it
constexpr int kRankMask = (1 << kRankBits) - 1;
constexpr int kScanPer = 4;
constexpr int kScanBlock = 1024;
constexpr int kScanTile = kScanBlock * kScanPer;  // 4096 entries per scan block
constexpr int kResolveThreads = 1024;
constexpr int kLdsSort = 2048;  // entries sorted in LDS by k_resolve

struct Layout {
  int nlist;
  int captot;               // list slots (sum of list capacities)
  int reccap;               // records per buffer
  int base[kMaxList + 1];   // first slot of each list
};

// One buffer of molecule records.
struct Recs {
  double *px, *py, *pz;
  float *ox, *oy, *oz;  // FP32-offset records: position - origin of `box` (null in FP64 mode)
  int* inst;      // simulation instance (ensembles only; null for one simulation)
  int* ident;     // species, 0 for a record whose molecule died
  uint64_t* serno;
  int* lpos;      // position in the reference's live list of its list
  int* box;       // virtual box (-1: not binned this step)
  int* rank;      // arrival rank within the box
  int* ov;        // surfaces only: order value of the record within its box list (k_relist)
  int* sb;        // surfaces only: 0 in solution, else (panel << 3) | bound state (1 front, 2 back)
};

struct Binned {
  int* count;   // [nlist*nbox*zsub+1] per simulation, zeroed by k_hscan after reading
  int* boff;    // sums of the 4096-key tiles
  int* bstart;  // first destination record of (list, box, z sub-bin); bstart[nscan-1] = records binned
  float4* spk;  // FP32 position + species bits of the destination records
};

struct Prod {
  int ident;
  int list;
  double x, y, z;
  int sb;  // surfaces: a product adsorbed by the reborn check (pending its list change), kProdKilled, else 0
  int pad;
  uint64_t serno;  // a confspread product keeps its reactant's serial number; 0 for a new serial
  double sx, sy, sz;  // surfaces: where the product was made (start of its reborn surface check)
};

// Per-species screening constants of the pair search, passed by value.
struct PairConst {
  uint32_t reactmask[kMaxSpecies];  // partner species with a bimolecular reaction
  float r2f[kMaxSpecies];           // (largest binding radius + FP32 margin)^2
};

// Per-species constants of the diffusion step and the first-order reaction
// coins of one simulation, gathered from its tables when the engine is made.
struct SpecRow {
  double ds;                   // diffusion step length (difstep)
  double prob[kMaxRxnPerSet];  // first-order reaction probabilities in table order; -1 for a
                               // reaction not permitted to solution-phase reactants (never fires)
  int ll;                      // list of the solution state
  int nu;                      // first-order reactions
  int rx[kMaxRxnPerSet];
};

// One simulation of the batch, as the kernels see it (device memory).
// host pointer -> the view's global pointer type (SMOLAMD_GLOBAL, engine.h)
#define GAS SMOLAMD_GLOBAL
#define AS_VIEW(T, p) ((GAS T*)(p))

struct Inst {
  Layout L;        // list slots in the batch-wide slot numbering
  Geometry g;
  DevSurf sf;
  GAS const DevTables* tab;
  GAS const SpecRow* srow;  // kMaxSpecies rows
  GAS Control* ctl;
  PairConst pc;
  uint32_t key0, key1;
  int keybase;     // first histogram key of this simulation
  // z sub-bins per virtual box (1, 2, 4 or 8): records are binned by (list,
  // box, z sub-bin), so a box's records are ordered by z slab; a sub-bin is at
  // least as tall as the largest screening radius, so partners within it lie
  // in the same or an adjacent sub-bin of a neighbor column
  int zsub;
  double screen_lim;  // coordinates beyond this bypass the FP32 pair screen
  GAS UniEvent* uniev;
  int maxevents;
  GAS Edge* edges;
  int maxedges;
  GAS int* estate;
  GAS int* accepted;
  GAS LogRow* log;
  GAS uint64_t* sk;
  GAS int* sv;
  int sortcap;
  GAS MoveEv* moves;  // list changes of the step (surfaces), capacity maxevents
  GAS int* killslot;
  GAS Prod* prods;
  // bounce by reflection: each source record's position before the step's
  // diffusion (3 per record) and, per destination record, its source record
  // (null when the model has no such reaction)
  GAS double* qpos;
  GAS int* src;
};

// Batch-wide counters.  The record count of the current buffer is kept per
// step parity: k_resolve of step s allocates product records with nprod[s&1]
// and raises nrec[s&1]; kernels of step s read nrec[(s+1)&1].
struct Glob {
  int nrec[2];
  int nprod[2];
  int ninst;  // simulations in the batch
  int halt;   // set on an inconsistency: every later step does nothing
  int ncand;  // screened candidate pairs of the step (k_pairs -> k_exact), reset by k_resolve
  int par;    // (s+1)&1 for the step s about to run: which nrec its kernels read
};

// kernels timed by engine_profile, each with its own start / stop events
enum TimedSlot : int { kTDiffuse, kTHsum, kTScan2, kTHscan, kTScatter, kTRelist, kTPairs, kTExact, kTResolve,
                       kTTile1, kTTileScan, kTTile2, kTimedSlots };

struct DeviceState {
  int ninst = 0;
  bool multi = false;                // more than one simulation: records carry their instance
  bool f32 = false;                  // FP32-offset records
  std::vector<EngineConfig> cfg;     // per simulation
  std::vector<DevTables> htab;       // host copies of the simulation tables
  std::vector<Inst> hinst;           // host copy of the instance views
  Inst* dinst = nullptr;
  Glob* glob = nullptr;
  int reccap = 0;                    // records per buffer
  int captot = 0;                    // list slots of all simulations
  double* gtab = nullptr;
  int ngauss = 0;
  Recs buf[2] = {};
  int cur = 0;                       // buffer holding the current state
  int* inv = nullptr;                // list slot -> record of the current buffer
  Binned bn = {};
  int nscan = 0, nscanblk = 0;
  uint64_t* bestkey = nullptr;
  int* killed = nullptr;
  // per-simulation arrays (the instance views point into them)
  Control* ctl = nullptr;
  DevTables* tab = nullptr;
  SpecRow* srow = nullptr;  // kMaxSpecies rows per simulation
  UniEvent* uniev = nullptr;
  Edge* edges = nullptr;
  int* estate = nullptr;
  int* accepted = nullptr;
  LogRow* log = nullptr;
  uint64_t* sortkey = nullptr;
  int* sortval = nullptr;
  int* killslot = nullptr;
  Prod* prods = nullptr;
  // surfaces (single simulation)
  bool surf_on = false;
  bool surf_curved = false;  // a triangle, cylinder or hemisphere panel
  DevPanel* d_pnl = nullptr;
  int* d_bpstart = nullptr;
  int* d_bplist = nullptr;
  DevSrfAct* d_act = nullptr;
  DevSrfAct* d_actb = nullptr;
  MoveEv* moves = nullptr;
  int* wq = nullptr;  // absorbing walls: survivors the second wall pass kills
  double* inj = nullptr;
  long injcap = 0;
  int hnrec = 0;                     // records uploaded so far
  // molecules added between steps (engine_append_fresh): first record, list
  // of each record since then, step they wait for
  int fresh_rec0 = 0;
  std::vector<int> fresh_list, fresh_lp;  // list and live position of each
  uint32_t fresh_step = 0;
  // tile binning (one simulation without surfaces); `sorted`: the current
  // buffer holds a binned step's records in key order with its box starts
  bool tile_ok = false;
  bool sorted = false;
  struct TileArrays {
    int ntile = 0, nkeys = 0, tshift = 0;
    int *stay = nullptr, *imm = nullptr, *toff = nullptr, *nkey = nullptr, *bucket = nullptr, *nspill = nullptr;
    int2 *range = nullptr, *spill = nullptr;
  } ta;
  double* qpos = nullptr;            // bounce by reflection: pre-diffusion positions, source-record order
  int* srcmap = nullptr;             // bounce by reflection: destination record -> source record
  int4* cand = nullptr;              // screened candidate pairs (k_pairs -> k_exact)
  int candcap = 0;
  bool half1 = false;                // every simulation reacts within one list only (half stencil)
  hipStream_t stream = nullptr;
  hipGraphExec_t graph[2] = {};
  int graph_ready[2] = {0, 0};
  hipEvent_t ev[16] = {};
  hipEvent_t kev[2 * kTimedSlots] = {};  // start / stop of each timed kernel (engine_profile)
  int timed_mask = 0;                    // slots the last timed step launched
  double last_ms = 0;
  int* obs = nullptr;                // observation rows (species counters between steps)
};

// ---------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ int pos2box(const Geometry& g, double x, double y, double z) {
  int b = 0;
  double p[3] = {x, y, z};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int indx = (int)((p[d] - g.bmin[d]) / g.bsize[d]);
    if (indx < 0) indx = 0;
    else if (indx >= g.side[d]) indx = g.side[d] - 1;
    b = g.side[d] * b + indx;
  }
  return b;
}

// Index of the bound state of a surface binding word: 0 front, 1 back.
__device__ __forceinline__ int bstate(int sb) { return (sb & 1) ? 0 : 1; }

// FP64 origin of virtual box b (the corner pos2box measures from); host and
// device evaluate the same expression, so FP32-offset records decode alike.
__host__ __device__ __forceinline__ void box_origin(const Geometry& g, int b, double* o) {
  const int iz = b % g.side[2], t = b / g.side[2];
  const int iy = t % g.side[1], ix = t / g.side[1];
  o[0] = g.bmin[0] + (double)ix * g.bsize[0];
  o[1] = g.bmin[1] + (double)iy * g.bsize[1];
  o[2] = g.bmin[2] + (double)iz * g.bsize[2];
}

// Record position: FP64 arrays, or FP32 offsets from the origin of box `bx`.
template <bool kF32>
__device__ __forceinline__ void load_pos(const Recs& a, const Geometry& g, int r, int bx, double* p) {
  if (kF32) {
    double o[3];
    box_origin(g, bx, o);
    p[0] = o[0] + (double)a.ox[r];
    p[1] = o[1] + (double)a.oy[r];
    p[2] = o[2] + (double)a.oz[r];
  } else {
    p[0] = a.px[r];
    p[1] = a.py[r];
    p[2] = a.pz[r];
  }
}

template <bool kF32>
__device__ __forceinline__ void store_pos(const Recs& a, const Geometry& g, int r, int bx, const double* p) {
  if (kF32) {
    double o[3];
    box_origin(g, bx, o);
    a.ox[r] = (float)(p[0] - o[0]);
    a.oy[r] = (float)(p[1] - o[1]);
    a.oz[r] = (float)(p[2] - o[2]);
  } else {
    a.px[r] = p[0];
    a.py[r] = p[1];
    a.pz[r] = p[2];
  }
}

}  // namespace smolamd
#include "surface_dev.h"
namespace smolamd {

// checkwalls for one molecule, walls in the reference's order 0..5.  Both
// outcomes of a wall are computed and one is selected, so the wall
// parameters load together and no lane waits on a branch.
__device__ __forceinline__ void apply_walls(const Geometry& g, double* p) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int wl = g.wall[2 * d], wh = g.wall[2 * d + 1];
    const double lo = g.low[d], hi = g.high[d];
    double v = p[d];
    const double vl = wl == kWallReflect ? 2 * lo - v : v + (hi - lo);
    if ((wl == kWallReflect || wl == kWallPeriodic) && v < lo) v = vl;
    const double vh = wh == kWallReflect ? 2 * hi - v : v + (lo - hi);
    if ((wh == kWallReflect || wh == kWallPeriodic) && v > hi) v = vh;
    p[d] = v;
  }
}

// checkwalls (smolwall.c:368-424) with absorbing walls, for one molecule:
// the walls in the reference's order (low then high side of each axis),
// reflective and periodic ones as apply_walls; an absorbing one kills a
// molecule beyond it, or one inside with the Brownian-bridge probability
// exp(-2 d_i d / step^2) that its path touched the wall (d_i, d: distances of
// the step's start and end from the wall).  Draws: Philox keyed by (serial,
// step, SMOLRNG_WALL, wall + 8 * pass), pass 1 for the wall pass after the
// reactions.  Returns 1 when killed (the later walls no longer matter).
__device__ __forceinline__ int walls_absorb(const Geometry& g, double* p, const double* px, double step,
                                            uint64_t ser, uint32_t stepidx, uint32_t k0, uint32_t k1, int pass) {
#pragma unroll
  for (int w = 0; w < 6; w++) {
    const int d = w >> 1, side = w & 1, t = g.wall[w];
    const double wp = side ? g.high[d] : g.low[d];
    if (t == kWallReflect) {
      if (side ? p[d] > wp : p[d] < wp) p[d] = 2 * wp - p[d];
    } else if (t == kWallPeriodic) {
      if (side ? p[d] > wp : p[d] < wp) p[d] += (side ? g.low[d] : g.high[d]) - wp;
    } else if (t == kWallAbsorb) {
      const double diff = wp - p[d], difi = wp - px[d];
      if ((!side && diff > 0) || (side && diff < 0)) return 1;
      const smolrng4 r = smolrng_mol(k0, k1, ser, stepidx, SMOLRNG_WALL, (uint32_t)(w + 8 * pass));
      if (smolrng_u01(r.v[0]) < exp(-2 * difi * diff / step / step)) return 1;
    }
  }
  return 0;
}

constexpr int kSbWallMark = 1 << 30;  // Recs::sb of a survivor the second wall pass kills (| its queue index)

__device__ __forceinline__ uint64_t atomic_load_u64(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int atomic_load_i32(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// k_diffuse: diffusion, walls, box, first-order reaction coins, histogram

// Histogram increment aggregated over runs of equal keys in a wavefront:
// records are stored in box order, so neighbouring lanes usually share a box
// counter; one atomic per run returns the run's base and each lane takes
// base + its offset in the run.  Every lane of the wave must call this.
__device__ __forceinline__ int wave_hist_rank(int* bcount, int key) {
  int lane = threadIdx.x & 63;
  int prev = __shfl_up(key, 1);
  bool start = (lane == 0) || (prev != key);
  unsigned long long starts = __ballot(start);
  unsigned long long le = starts & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
  int leader = 63 - __clzll(le);
  unsigned long long above = (leader == 63) ? 0ull : (starts & ~((2ull << leader) - 1));
  int next = above ? (__ffsll((long long)above) - 1) : 64;
  int base = 0;
  if (lane == leader && key >= 0) base = atomicAdd(&bcount[key], next - leader);
  base = __shfl(base, leader);
  return base + (lane - leader);
}

// Record count of the current buffer (see Glob), read by the kernels of a
// step before k_resolve.
__device__ __forceinline__ int cur_nrec(const Inst* insts, const Glob* gl) {
  (void)insts;
  // both counts and the parity load together (one scalar load, no chain)
  const int n0 = gl->nrec[0], n1 = gl->nrec[1];
  return gl->par ? n1 : n0;
}

// The same count, -1 once stepping has halted: the halt flag, both counts and
// the parity come in one scalar load at the start of a pass.
__device__ __forceinline__ int live_nrec(const Glob* gl) {
  const int n0 = gl->nrec[0], n1 = gl->nrec[1], h = gl->halt, p = gl->par;
  return h ? -1 : (p ? n1 : n0);
}

// Wavefront-private copies of the species rows and of each list's diffusion
// limit (list positions below it diffuse this step; INT_MIN for a list that
// does not diffuse), for single-simulation passes: each lane loads one row
// or one list's counters while the wave's record loads are in flight, so no
// per-molecule table lookup waits on a dependent global load.  Every lane of
// the wave must call this.
__device__ __forceinline__ void fill_spec_rows(const Inst& I, SpecRow* S, int* lim) {
  const int lane = threadIdx.x & 63;
  const Control* ctl = I.ctl;
  const int ls = lane & (kMaxList - 1);
  const SpecRow row = I.srow[lane & (kMaxSpecies - 1)];
  const int dl = I.tab->diffuselist[ls], nl = ctl->nl[ls], nf = ctl->nfresh[ls];
  if (lane < kMaxSpecies) S[lane] = row;
  if (lane < kMaxList) lim[lane] = dl ? nl - nf : INT_MIN;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int kSurf, bool kMulti, bool kF32>
__global__ __launch_bounds__(kBlock) void k_diffuse(const Inst* __restrict__ insts, const Glob* __restrict__ gl,
                                                    Recs a, int* __restrict__ bcount,
                                                    const double* __restrict__ gtab, int gmask,
                                                    const double* __restrict__ inj, int* __restrict__ inv) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int hkey = -1;  // histogram key (simulation's first key + (list*nbox + box)*zsub + sub-bin), -1 when not binned
  int zsbits = 0;  // z sub-bin, kept above the rank for k_scatter
  // uniform over the block: a block without records (or any block once
  // stepping has halted) has nothing to bin
  const int n = live_nrec(gl);
  if ((int)(blockIdx.x * blockDim.x) >= n) return;
  // the record's fields, loaded by every lane (a lane past the last record
  // reads the last one and discards it) before anything waits on them
  const int rr = r < n ? r : n - 1;
  const int i0 = a.ident[rr];
  const int in0 = kMulti ? a.inst[rr] : 0;
  const int sb0 = kSurf ? a.sb[rr] : 0;
  const int lp0 = a.lpos[rr];
  const uint64_t ser0 = a.serno[rr];
  const int box0 = (kF32 || kSurf) ? a.box[rr] : 0;
  double pos0[3];
  load_pos<kF32>(a, insts[(in0 >= 0 && in0 < gl->ninst) ? in0 : 0].g, rr, box0, pos0);
  __shared__ SpecRow s_spec[kMulti ? 1 : kBlock / 64][kMaxSpecies];
  __shared__ int s_lim[kMulti ? 1 : kBlock / 64][kMaxList];
  SpecRow* S = s_spec[kMulti ? 0 : threadIdx.x >> 6];
  int* Lim = s_lim[kMulti ? 0 : threadIdx.x >> 6];
  const uint32_t step0 = kMulti ? 0u : insts[0].ctl->step;
  if (!kMulti) fill_spec_rows(insts[0], S, Lim);
  if (r < n) {
    int i = i0;
    int in = in0;
    if (i <= 0 || i >= kMaxSpecies || in < 0 || in >= gl->ninst) {
      if (i != 0) atomicOr((int*)&gl->halt, 1);
      a.box[r] = -1;
    } else {
      const Inst& I = insts[in];
      const DevTables* tab = I.tab;
      Control* ctl = I.ctl;
      const Geometry& g = I.g;
      int ll = kMulti ? tab->listof[i] : S[i].ll;
      int sb = 0;  // surface binding (kSurf): bound molecules live in the list of their state
      if (kSurf) {
        sb = sb0;
        if (sb) ll = I.sf.blist[i][bstate(sb)];
      }
      const int lp = lp0;
      const uint64_t ser = ser0;
      const uint32_t step = kMulti ? ctl->step : step0;
      int mkind = 0, newll = ll;  // list change queued for the first molsort (surfaces)
      int obox = -1, rbox = -1;    // desorption: box in the old list, box the reference files it in
      const int bold = kF32 ? box0 : 0;
      const int bprev = kSurf ? box0 : -1;  // the box the record was filed in last step
      double pos[3] = {pos0[0], pos0[1], pos0[2]};
      if (!kSurf && !kMulti && I.qpos) {  // bounce by reflection: the position before diffusion (posx)
        I.qpos[3 * (long)r] = pos[0];
        I.qpos[3 * (long)r + 1] = pos[1];
        I.qpos[3 * (long)r + 2] = pos[2];
      }
      // molecules added since the last step sit at the end of their list and
      // do not diffuse in this step
      const bool moved =
          kMulti ? (tab->diffuselist[ll] && lp < ctl->nl[ll] - ctl->nfresh[ll]) : lp < Lim[ll];
      bool flipped = false;  // a bound molecule moved to the other face of its panel
      bool dead = false;     // absorbed by a surface: not binned, its slot dies in the first molsort
      bool wall2 = false;    // absorbing walls: killed by the second wall pass unless it reacts
      if (moved) {
        const double posx[3] = {pos[0], pos[1], pos[2]};
        double ds = kMulti ? tab->difstep[i] : S[i].ds;
        double gg[3];
        if (inj) {
          long k = (long)ctl->pfx[ll] + lp;
          gg[0] = inj[3 * k];
          gg[1] = inj[3 * k + 1];
          gg[2] = inj[3 * k + 2];
        } else {
          smolrng4 rn = smolrng_mol(I.key0, I.key1, ser, step, SMOLRNG_DIFFUSE, 0);
          gg[0] = gtab[rn.v[0] & (uint32_t)gmask];
          gg[1] = gtab[rn.v[1] & (uint32_t)gmask];
          gg[2] = gtab[rn.v[2] & (uint32_t)gmask];
        }
        pos[0] += ds * gg[0];
        pos[1] += ds * gg[1];
        pos[2] += ds * gg[2];
        if (!kSurf) {
          apply_walls(g, pos);
        } else if (I.sf.walls) {
          // no surfaces: checkwalls with absorbing walls; a killed molecule
          // leaves its list in the first molsort
          if (walls_absorb(g, pos, posx, ds, ser, step, I.key0, I.key1, 0)) {
            dead = true;
            mkind = kMoveKill;
            atomicSub(&ctl->count[i], 1);
          } else {
            // the wall pass after the reactions (smolsim.cpp:2660-2666)
            // flips the absorbing coins of every molecule again; one that
            // does not react is still where this pass left it, so its coin
            // is drawn here and the kill made at the end of the step
            double p2[3] = {pos[0], pos[1], pos[2]};
            wall2 = walls_absorb(g, p2, posx, ds, ser, step, I.key0, I.key1, 1);
            if (p2[0] != pos[0] || p2[1] != pos[1] || p2[2] != pos[2]) {
              // still beyond a wall after the first pass: not run here
              atomicOr(&ctl->err, kErrSurface);
              atomicOr((int*)&gl->halt, 1);
            }
          }
        } else if (I.sf.listdiffuse[ll]) {
          // with surfaces the walls are not applied (smolsim.cpp:2618-2630)
          int i2 = i, bnd = 0;
          int st = srf_check<(kSurf > 1)>(I.sf, g, tab->difc, pos, posx, &i2, ser, step, SMOLRNG_SURF_DIFFUSED, I.key0, I.key1,
                             &bnd);
          if (st == kSrfAdsorbed && I.sf.blist[i2][bstate(bnd)] >= 0) {  // bound: to the list of its state
            sb = bnd;
            a.sb[r] = bnd;
            newll = I.sf.blist[i2][bstate(bnd)];
            mkind = kMoveList;
            atomicAdd((bnd & 3) == 1 ? &ctl->countf[i2] : &ctl->countk[i2], 1);
          } else if (st == kSrfAbsorbed) {
            // molkill (smolsurface.c:4994-4996): the first molsort removes
            // the slot; the molecule counts no more, under any species
            dead = true;
            mkind = kMoveKill;
            atomicSub(&ctl->count[i], 1);
            i2 = i;
          } else if (st != kSrfOk || (i2 != i && tab->listof[i2] != ll)) {
            // an outcome the device path does not run (adsorption into a
            // state without a list, a species change into another list):
            // report it and stop stepping
            atomicOr(&ctl->err, kErrSurface);
            atomicOr((int*)&gl->halt, 1);
          }
          if (i2 != i) {
            a.ident[r] = i2;
            atomicSub(&ctl->count[i], 1);
            atomicAdd(&ctl->count[i2], 1);
            i = i2;
          }
        }
      } else if (kSurf && sb && I.sf.listbound[ll]) {
        // checksurfacebound (smolsurface.c:5184-5199): desorption, a flip or
        // absorption.  The molecule leaves the box it was filed in (bound
        // lists are not reassigned, so that box may be stale).
        int i2 = i, rb = 0;
        obox = bprev >= 0 ? bprev : pos2box(g, pos[0], pos[1], pos[2]);
        const int st = srf_bound<(kSurf > 1)>(I.sf, g, tab->difc, tab->difstep, pos, &i2, sb, ser, step, I.key0, I.key1,
                                              &rbox, &rb);
        if (st == kSrfDesorbed && (!rb || I.sf.blist[i2][bstate(rb)] >= 0)) {
          atomicSub((sb & 3) == 1 ? &ctl->countf[i] : &ctl->countk[i], 1);
          if (i2 != i) {
            a.ident[r] = i2;
            atomicSub(&ctl->count[i], 1);
            atomicAdd(&ctl->count[i2], 1);
            i = i2;
          }
          // the resurrected molecule joins the list of its state: solution,
          // or the bound state its surface check adsorbed it to
          sb = rb;
          a.sb[r] = rb;
          newll = rb ? I.sf.blist[i][bstate(rb)] : tab->listof[i];
          if (rb) atomicAdd((rb & 3) == 1 ? &ctl->countf[i] : &ctl->countk[i], 1);
          mkind = kMoveDesorb;
        } else if (st == kSrfFlipped && I.sf.blist[i2][bstate(rb)] >= 0) {
          // a flip: the state's counter, the species and, when the new
          // state lives in another list, a list change for the first molsort
          atomicSub((sb & 3) == 1 ? &ctl->countf[i] : &ctl->countk[i], 1);
          atomicAdd((rb & 3) == 1 ? &ctl->countf[i2] : &ctl->countk[i2], 1);
          if (i2 != i) {
            a.ident[r] = i2;
            atomicSub(&ctl->count[i], 1);
            atomicAdd(&ctl->count[i2], 1);
            i = i2;
          }
          sb = rb;
          a.sb[r] = rb;
          flipped = true;
          if (I.sf.blist[i][bstate(rb)] != ll) {
            newll = I.sf.blist[i][bstate(rb)];
            mkind = kMoveList;
          }
        } else if (st == kSrfAbsorbed) {
          dead = true;
          mkind = kMoveKill;
          atomicSub((sb & 3) == 1 ? &ctl->countf[i] : &ctl->countk[i], 1);
          atomicSub(&ctl->count[i], 1);
        } else if (st != kSrfOk) {
          atomicOr(&ctl->err, kErrSurface);
          atomicOr((int*)&gl->halt, 1);
        }
      }
      int mev = -1;
      if (kSurf && mkind) {  // queue the list change; the first molsort gives the record its new slot
        const int e = atomicAdd(&ctl->nmove, 1);
        if (e < I.maxevents) {
          MoveEv mv;
          mv.rec = r;
          mv.dst = -1;
          mv.slot = I.L.base[ll] + lp;
          mv.list = newll;
          mv.kind = mkind;
          mv.obox = obox;
          mv.nbox = -1;
          mv.ov = lp;
          I.moves[e] = mv;
          mev = e;
          a.lpos[r] = -(e + 2);
        } else {  // a lost list change would corrupt the lists: stop
          atomicOr(&ctl->err, kErrEventOverflow);
          atomicOr((int*)&gl->halt, 1);
        }
        ll = newll;
      }
      int b = pos2box(g, pos[0], pos[1], pos[2]);
      if (kSurf && mkind == kMoveDesorb) b = rbox;  // the reference's (possibly stale) box for this step
      // a bound molecule in a list that does not diffuse keeps the box it was
      // filed in: reassignmolecs skips such lists (smolboxes.c:896-954), so a
      // flip to the other face of a panel on a box boundary leaves it stale
      else if (kSurf && sb && !moved && bprev >= 0) b = bprev;
      if (kSurf && mev >= 0) {  // boxes of the list change: a moving molecule keeps its box
        // a molecule of a diffusing list was filed in the box of its new
        // position by reassignmolecs before the molsort
        if (mkind == kMoveList || (mkind == kMoveKill && moved)) I.moves[mev].obox = b;
        I.moves[mev].nbox = b;
      }
      // FP32-offset records are stored relative to the new box
      if (!dead && (moved || flipped || mkind == kMoveDesorb || (kF32 && b != bold))) store_pos<kF32>(a, g, r, b, pos);
      // z sub-bin: floor of the box-relative z coordinate times zsub (exact
      // for a power of two), clamped like the box index
      int zs = 0;
      if (!dead && I.zsub > 1) {
        const double zrel = (pos[2] - g.bmin[2]) / g.bsize[2];
        int iz = (int)zrel;
        if (iz < 0) iz = 0;
        else if (iz >= g.side[2]) iz = g.side[2] - 1;
        zs = (int)((zrel - (double)iz) * (double)I.zsub);
        if (zs < 0) zs = 0;
        else if (zs >= I.zsub) zs = I.zsub - 1;
      }
      hkey = dead ? -1 : I.keybase + (ll * g.nbox + b) * I.zsub + zs;
      zsbits = zs << kRankBits;
      // unireact: coin flips in table order, first success reacts; a
      // molecule that reacts takes no part in the pair search, so it is not
      // binned
      int nu = (sb || dead) ? 0 : (kMulti ? tab->nuni[i] : S[i].nu);  // surface-bound molecules do not react
      if (nu > 0) {
        smolrng4 rb;
        int blk = -1;
        for (int j = 0; j < nu; j++) {
          int rx;
          double prob;
          if (kMulti) {
            rx = tab->uni[i][j];
            const DevRxn& d = tab->rxn1[rx];
            if (!d.permit_soln) continue;
            prob = d.prob;
          } else {
            rx = S[i].rx[j];
            prob = S[i].prob[j];
          }
          if (blk != j / 4) {
            blk = j / 4;
            rb = smolrng_mol(I.key0, I.key1, ser, step, SMOLRNG_UNIRXN, (uint32_t)blk);
          }
          if (smolrng_u01(rb.v[j & 3]) < prob) {
            int e = atomicAdd(&ctl->nuni, 1);
            if (e < I.maxevents) {
              I.uniev[e].rec = r;
              I.uniev[e].slot = mkind ? -1 : I.L.base[ll] + lp;  // -1: appended by this step's first molsort
              I.uniev[e].r = rx;
              I.uniev[e].box = b;
            } else {
              atomicOr(&ctl->err, kErrEventOverflow);
            }
            hkey = -1;
            // the first molsort may move this molecule's slot: it reads the
            // record of an unbinned slot from the inverse map
            if (kSurf && !mkind && lp >= ctl->nl[ll] - I.sortcap) inv[I.L.base[ll] + lp] = -(r + 2);
            break;
          }
        }
      }
      if (kSurf && wall2 && hkey >= 0) {  // (a first-order reactant dies anyway)
        const int q = atomicAdd(&ctl->nwall, 1);
        if (q < I.maxevents) {
          I.sf.wq[q] = -1;
          a.sb[r] = kSbWallMark | q;
        } else {
          atomicOr(&ctl->err, kErrEventOverflow);
          atomicOr((int*)&gl->halt, 1);
        }
      }
      a.box[r] = hkey < 0 ? -1 : b;
    }
  }
  int rank = wave_hist_rank(bcount, hkey);
  if (hkey >= 0) a.rank[r] = rank | zsbits;
}

// ---------------------------------------------------------------------------
// histogram scan

// exclusive scan of block sums, one workgroup (up to 1024*8 tiles)
__global__ __launch_bounds__(kScanBlock) void k_scan2(int* __restrict__ bsum, int nblk) {
  __shared__ int s[kScanBlock];
  constexpr int per = 8;
  int v[per];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < per; k++) {
    int i = threadIdx.x * per + k;
    v[k] = i < nblk ? bsum[i] : 0;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int x = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  int run = s[threadIdx.x] - sum;
#pragma unroll
  for (int k = 0; k < per; k++) {
    int i = threadIdx.x * per + k;
    if (i < nblk) bsum[i] = run;
    run += v[k];
  }
}

// ---------------------------------------------------------------------------
// Histogram scan in two passes over 4096-key tiles (256 threads x 16 keys):
// k_hsum writes each tile's sum; k_hscan adds the sums of the tiles before
// its own (or reads them already scanned by k_scan2 when there are many
// tiles), scans its tile with wave shuffles and writes the box starts,
// zeroing the counts for the next step.
constexpr int kHThreads = 256;
constexpr int kHPer = 16;
constexpr int kHTile = kHThreads * kHPer;  // keys per tile
constexpr int kHFusedTiles = 1024;         // up to this many tiles k_hscan sums the tile sums itself

__device__ __forceinline__ void load16(const int* __restrict__ p, int i0, int n, int* v) {
  if (i0 + kHPer <= n) {
    const int4* q = reinterpret_cast<const int4*>(p + i0);
#pragma unroll
    for (int k = 0; k < kHPer / 4; k++) {
      const int4 w = q[k];
      v[4 * k] = w.x;
      v[4 * k + 1] = w.y;
      v[4 * k + 2] = w.z;
      v[4 * k + 3] = w.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kHPer; k++) v[k] = (i0 + k < n) ? p[i0 + k] : 0;
  }
}

// inclusive scan of one value per lane within a wavefront
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

// block-wide sum of one value per thread (kHThreads threads)
__device__ __forceinline__ int block_sum(int x, int* s_w) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = x;
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int w = 0; w < kHThreads / 64; w++) t += s_w[w];
  return t;
}

__global__ __launch_bounds__(kHThreads) void k_hsum(const int* __restrict__ count, int* __restrict__ bsum, int n) {
  __shared__ int s_w[kHThreads / 64];
  int v[kHPer];
  load16(count, blockIdx.x * kHTile + threadIdx.x * kHPer, n, v);
  int t = 0;
#pragma unroll
  for (int k = 0; k < kHPer; k++) t += v[k];
  t = block_sum(t, s_w);
  if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}

// scanned != 0: bsum holds each tile's exclusive offset (k_scan2)
__global__ __launch_bounds__(kHThreads) void k_hscan(int* __restrict__ count, const int* __restrict__ bsum,
                                                     int* __restrict__ bstart, int n, int scanned) {
  __shared__ int s_w[kHThreads / 64];
  __shared__ int s_x[kHThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i0 = blockIdx.x * kHTile + tid * kHPer;
  int v[kHPer];
  load16(count, i0, n, v);
  // the tile's offset
  int off;
  if (scanned) {
    off = bsum[blockIdx.x];
  } else {
    int part = 0;
    for (int k = tid; k < (int)blockIdx.x; k += kHThreads) part += bsum[k];
    off = block_sum(part, s_w);
  }
  // thread-local exclusive prefix, then the thread's offset in the tile
  int t = 0;
#pragma unroll
  for (int k = 0; k < kHPer; k++) {
    const int c = v[k];
    v[k] = t;
    t += c;
  }
  int inc = t;  // inclusive scan of the thread totals within the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) s_x[wv] = inc;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wv; w++) wbase += s_x[w];
  const int base = off + wbase + inc - t;
  if (i0 + kHPer <= n) {
    int4* q = reinterpret_cast<int4*>(bstart + i0);
    int4* z = reinterpret_cast<int4*>(count + i0);
#pragma unroll
    for (int k = 0; k < kHPer / 4; k++) {
      q[k] = make_int4(base + v[4 * k], base + v[4 * k + 1], base + v[4 * k + 2], base + v[4 * k + 3]);
      z[k] = make_int4(0, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kHPer; k++)
      if (i0 + k < n) {
        bstart[i0 + k] = base + v[k];
        count[i0 + k] = 0;
      }
  }
}

// counting-sort scatter of the binned records into the destination buffer
// The inverse map (list slot -> record) is written only for the last
// `window` slots of each list: compaction in k_resolve reads tail slots only,
// and the kills of one step never exceed the window.
template <bool kMulti, bool kF32, int kSurf>
__global__ __launch_bounds__(kBlock) void k_scatter(const Inst* __restrict__ insts, const Glob* __restrict__ gl,
                                                    Recs a, Recs b, Binned bn, int* __restrict__ inv) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (gl->halt || r >= cur_nrec(insts, gl)) return;
  int bx = a.box[r];
  if (bx < 0) return;
  int in = kMulti ? a.inst[r] : 0;
  const Inst& I = insts[in];
  int i = a.ident[r];
  int ll = I.tab->listof[i];
  int sb = kSurf ? a.sb[r] : 0;
  int wmark = -1;  // queue index of a survivor the second wall pass kills
  if (kSurf && (sb & kSbWallMark)) {
    wmark = sb & (kSbWallMark - 1);
    sb = 0;
  }
  if (kSurf && sb) ll = I.sf.blist[i][bstate(sb)];
  const int rk = a.rank[r];
  const int zs = rk >> kRankBits;
  const int key = I.keybase + (ll * I.g.nbox + bx) * I.zsub + zs;
  int dst = bn.bstart[key] + (rk & kRankMask);
  // the key a record was binned with, read by the tile binning of the next
  // step (one simulation without surfaces), and for bounce by reflection the
  // record's source record
  if (!kMulti && !kSurf) {
    b.rank[dst] = key;
    if (I.src) I.src[dst] = r;
  }
  double x, y, z;
  if (kF32) {
    const float fx = a.ox[r], fy = a.oy[r], fz = a.oz[r];
    b.ox[dst] = fx;
    b.oy[dst] = fy;
    b.oz[dst] = fz;
    double o[3];
    box_origin(I.g, bx, o);
    x = o[0] + (double)fx;
    y = o[1] + (double)fy;
    z = o[2] + (double)fz;
  } else {
    x = a.px[r];
    y = a.py[r];
    z = a.pz[r];
    b.px[dst] = x;
    b.py[dst] = y;
    b.pz[dst] = z;
  }
  int lp = a.lpos[r];
  if (kMulti) b.inst[dst] = in;
  b.ident[dst] = i;
  b.serno[dst] = a.serno[r];
  b.lpos[dst] = lp;
  b.box[dst] = bx;
  // bit 16 of the payload word marks a coordinate beyond the screen's bound:
  // such a molecule bypasses the FP32 screen
  const double lim = I.screen_lim;
  int far = !(fabs(x) <= lim && fabs(y) <= lim && fabs(z) <= lim);
  // bit 20: a surface-bound molecule, which takes no part in the pair search
  bn.spk[dst] = make_float4((float)x, (float)y, (float)z,
                            __int_as_float(i | (far << 16) | (zs << 17) | ((sb != 0) << 20)));
  if (kSurf) {
    b.sb[dst] = sb;
    if (wmark >= 0) I.sf.wq[wmark] = dst;
    b.ov[dst] = lp < 0 ? -1 : lp;  // box lists are in live order until the first molsort changes them
    if (lp < 0) {  // list change of this step: the first molsort gives the slot
      I.moves[-lp - 2].dst = dst;
      return;
    }
  }
  if (lp >= I.ctl->nl[ll] - I.sortcap) inv[I.L.base[ll] + lp] = dst;
}

// ---------------------------------------------------------------------------
// Tile binning (one simulation without surfaces): the diffusion stage and the
// counting sort into (list, box, z sub-bin) order in two launches.
//
// The key space is cut into tiles of 2^tshift consecutive keys.  The source
// buffer holds the last step's binned records in key order, so each tile's
// records are one range of it (read from the last step's box starts),
// followed by an unsorted tail: the last step's products and molecules added
// since.  Most records keep their tile from one step to the next; the few
// that leave it (a box change across the tile's edge) and the tail are
// queued for the tile of their new key.
//   k_tile1      advances every record (diffuse, checkwalls, pos2box, z
//                sub-bin, unireact coins; first-order events are queued
//                here), one thread per record; stores its new position and
//                key in the source buffer, counts the records that keep their
//                tile and queues the others;
//   k_tile_scan  one workgroup: each tile's first destination record (the
//                records of the tiles before it) and its source range;
//   k_tile2      reads the keys of the tile's records and of its queued
//                immigrants, ranks them per key in LDS, scans the tile,
//                writes the tile's box starts and copies the records to
//                their places.
// A record's `rank` field holds the key it was binned with (written by
// k_tile2 or k_scatter), from which k_tile1 knows its tile; k_tile1 leaves
// the new key in a separate array, so the old keys stay readable to the
// neighbors that compare with them.  The stay counts are added per run of records of one
// tile within a wavefront (records are in key order), so the global atomics
// of a step number about one per wavefront and hardly ever meet on one
// address.
constexpr int kTileMinShift = 8;   // at least kBlock keys per tile (the scan takes whole keys per thread)
constexpr int kTileMaxShift = 11;
constexpr int kTileMaxKeys = 1 << kTileMaxShift;
constexpr int kTileBucket = 64;     // queued immigrants per tile; more go to the shared spill list
constexpr int kTileCache = 4;       // records per thread that k_tile2 keeps in registers between its passes

struct Tiles {
  int ntile, nkeys, tshift, reccap;
  int* stay;    // [2 * ntile] records that keep their tile, by step parity (k_tile2 clears the other parity)
  int* imm;     // [2 * ntile] immigrants per tile, by step parity
  int* toff;    // [ntile] first destination record of each tile
  int2* range;  // [2 * ntile] source records [x, y) of the tile by step parity ((0, 0): none), set by k_tile1
  int* nkey;    // [record capacity] new key of each source record (-1: not binned), set by k_tile1
  int* bucket;  // [ntile * kTileBucket] source records of the immigrants
  int2* spill;  // [record capacity] (record, key) of immigrants beyond a full bucket
  int* nspill;  // [2] by step parity
};

// The fields of one source record that the diffusion stage reads.
struct TileIn {
  double pos[3];
  uint64_t ser;
  int i, lp, bold, okey;
};

template <bool kF32>
__device__ __forceinline__ void tile_fetch(const Recs& a, const Geometry& g, int r, TileIn& in) {
  in.i = a.ident[r];
  in.lp = a.lpos[r];
  in.ser = a.serno[r];
  in.okey = a.rank[r];  // meaningful only for a record binned by the last step
  in.bold = kF32 ? a.box[r] : 0;
  load_pos<kF32>(a, g, r, in.bold, in.pos);
}

// The diffusion stage of one record in three phases, so that a thread's
// records wait on their table loads together: tile_draw (the Philox draw of
// the displacement, smolmolec.c:3115-3168 diffuse, as table indices),
// tile_table (the Gaussian table loads, or the injected deviates) and
// tile_finish (the step, checkwalls smolwall.c:368-424, pos2box
// smolboxes.c:67-79, the z sub-bin and the unireact coins smolreact.c:
// 2997-3054), as k_diffuse computes them for a simulation without surfaces.
struct TileDraw {
  uint32_t gi[3];  // table indices of the three deviates
  int ll;          // list, -1: no molecule
  bool moved;      // diffuses this step
};

__device__ __forceinline__ void tile_draw(const Inst& I, const SpecRow* S, const int* Lim, const TileIn& in,
                                          uint32_t step, bool inj, int gmask, TileDraw& d) {
  const int i = in.i;
  d.ll = (i > 0 && i < kMaxSpecies) ? S[i].ll : -1;
  d.moved = d.ll >= 0 && in.lp < Lim[d.ll];
  d.gi[0] = d.gi[1] = d.gi[2] = 0;
  if (d.moved && !inj) {
    const smolrng4 rn = smolrng_mol(I.key0, I.key1, in.ser, step, SMOLRNG_DIFFUSE, 0);
    d.gi[0] = rn.v[0] & (uint32_t)gmask;
    d.gi[1] = rn.v[1] & (uint32_t)gmask;
    d.gi[2] = rn.v[2] & (uint32_t)gmask;
  }
}

__device__ __forceinline__ void tile_table(const Inst& I, const TileIn& in, const TileDraw& d,
                                           const double* __restrict__ gtab, const double* __restrict__ inj,
                                           double* gg) {
  gg[0] = gg[1] = gg[2] = 0.0;
  if (!d.moved) return;
  if (inj) {
    const long k = (long)I.ctl->pfx[d.ll] + in.lp;
    gg[0] = inj[3 * k];
    gg[1] = inj[3 * k + 1];
    gg[2] = inj[3 * k + 2];
  } else {
    gg[0] = gtab[d.gi[0]];
    gg[1] = gtab[d.gi[1]];
    gg[2] = gtab[d.gi[2]];
  }
}

// The new position (and, for FP32-offset records, the box it is stored
// relative to) goes back to the record; the return value is the record's new
// key, -1 for a record that is not binned (no molecule, or a first-order
// reactant, whose event is queued here).
template <bool kF32>
__device__ __forceinline__ int tile_finish(const Inst& I, const SpecRow* S, const Recs& a, int r, const TileIn& in,
                                           const TileDraw& d, const double* gg, uint32_t step, Glob* gl) {
  const int i = in.i, lp = in.lp, bold = in.bold, ll = d.ll;
  const uint64_t ser = in.ser;
  const Geometry& g = I.g;
  if (ll < 0) {
    if (i != 0) atomicOr((int*)&gl->halt, 1);
    return -1;
  }
  double pos[3] = {in.pos[0], in.pos[1], in.pos[2]};
  if (I.qpos) {  // the position before the step's diffusion (posx)
    I.qpos[3 * (long)r] = pos[0];
    I.qpos[3 * (long)r + 1] = pos[1];
    I.qpos[3 * (long)r + 2] = pos[2];
  }
  if (d.moved) {
    const double ds = S[i].ds;
    pos[0] += ds * gg[0];
    pos[1] += ds * gg[1];
    pos[2] += ds * gg[2];
    apply_walls(g, pos);
  }
  const int b = pos2box(g, pos[0], pos[1], pos[2]);
  // FP32-offset records are stored relative to the new box
  if (d.moved || (kF32 && b != bold)) {
    store_pos<kF32>(a, g, r, b, pos);
    if (kF32) a.box[r] = b;
  }
  int zs = 0;
  if (I.zsub > 1) {
    const double zrel = (pos[2] - g.bmin[2]) / g.bsize[2];
    int iz = (int)zrel;
    if (iz < 0) iz = 0;
    else if (iz >= g.side[2]) iz = g.side[2] - 1;
    zs = (int)((zrel - (double)iz) * (double)I.zsub);
    if (zs < 0) zs = 0;
    else if (zs >= I.zsub) zs = I.zsub - 1;
  }
  const int nu = S[i].nu;
  if (nu > 0) {
    smolrng4 rb;
    int blk = -1;
    for (int j = 0; j < nu; j++) {
      if (blk != j / 4) {
        blk = j / 4;
        rb = smolrng_mol(I.key0, I.key1, ser, step, SMOLRNG_UNIRXN, (uint32_t)blk);
      }
      if (smolrng_u01(rb.v[j & 3]) < S[i].prob[j]) {
        Control* ctl = I.ctl;
        const int e = atomicAdd(&ctl->nuni, 1);
        if (e < I.maxevents) {
          I.uniev[e].rec = r;
          I.uniev[e].slot = I.L.base[ll] + lp;
          I.uniev[e].r = S[i].rx[j];
          I.uniev[e].box = b;
        } else {
          atomicOr(&ctl->err, kErrEventOverflow);
        }
        return -1;
      }
    }
  }
  return I.keybase + (ll * g.nbox + b) * I.zsub + zs;
}

// Queue record r for the tile of its new key.
__device__ __forceinline__ void tile_queue(const Tiles& tl, int* imm, int* nspill, int key, int r) {
  const int dt = key >> tl.tshift;
  const int q = atomicAdd(&imm[dt], 1);
  if (q < kTileBucket) {
    tl.bucket[dt * kTileBucket + q] = r;
  } else {  // every record is queued at most once a step: the list never fills
    const int j = atomicAdd(nspill, 1);
    tl.spill[j] = make_int2(r, key);
  }
}

// Adds, for each run of lanes of one tile (tile >= 0) in the wavefront, the
// number of its lanes with flag set to cnt[tile].  Every lane of the wave
// must call this.
__device__ __forceinline__ void wave_run_count(int* cnt, int tile, bool flag) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(tile, 1);
  const bool start = lane == 0 || prev != tile;
  const unsigned long long starts = __ballot(start), fl = __ballot(flag);
  if (start && tile >= 0) {
    const unsigned long long above = starts & ~((2ull << lane) - 1);
    const int nxt = above ? __ffsll((long long)above) - 1 : 64;
    const unsigned long long run = (nxt == 64 ? ~0ull : ((1ull << nxt) - 1)) & ~((1ull << lane) - 1);
    const int c = __popcll(fl & run);
    if (c) atomicAdd(&cnt[tile], c);
  }
}

// kTile1Per records per thread, a block's records in kTile1Per runs of
// kBlock.  Every record load of the block goes out before anything waits
// (before the record count is known: lanes past the end read records that
// exist but are not used), then the draws of all of a thread's records, their
// table loads together, and the rest of the stage.
constexpr int kTile1Per = 2;
template <bool kF32>
__global__ __launch_bounds__(kBlock) void k_tile1(const Inst* __restrict__ insts, Glob* __restrict__ gl, Recs a,
                                                  const int* __restrict__ bstart, Tiles tl,
                                                  const double* __restrict__ gtab, int gmask,
                                                  const double* __restrict__ inj) {
  const Inst& I = insts[0];
  __shared__ SpecRow s_spec[kBlock / 64][kMaxSpecies];
  __shared__ int s_lim[kBlock / 64][kMaxList];
  SpecRow* S = s_spec[threadIdx.x >> 6];
  int* Lim = s_lim[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * kBlock * kTile1Per;
  const int rlast = tl.reccap - 1;
  TileIn in[kTile1Per];
  int nbk[kTile1Per];
#pragma unroll
  for (int j = 0; j < kTile1Per; j++) {
    const int r = min(r0 + j * kBlock + (int)threadIdx.x, rlast);
    tile_fetch<kF32>(a, I.g, r, in[j]);
    // the old keys of the records just outside the wave (first and last lane)
    nbk[j] = -1;
    if (lane == 0) nbk[j] = a.rank[r > 0 ? r - 1 : 0];
    if (lane == 63) nbk[j] = a.rank[min(r + 1, rlast)];
  }
  const int nb0 = bstart[tl.nkeys];  // records binned by the last step; the tail follows
  const int n = live_nrec(gl);
  if (n < 0 || r0 >= n) return;  // uniform over the block
  fill_spec_rows(I, S, Lim);
  const uint32_t step = I.ctl->step;
  const int par = step & 1;
  TileDraw d[kTile1Per];
#pragma unroll
  for (int j = 0; j < kTile1Per; j++) {
    tile_draw(I, S, Lim, in[j], step, inj != nullptr, gmask, d[j]);
    if (r0 + j * kBlock + (int)threadIdx.x >= n) d[j].ll = -2;  // past the last record
  }
  double gg[kTile1Per][3];
#pragma unroll
  for (int j = 0; j < kTile1Per; j++) tile_table(I, in[j], d[j], gtab, inj, gg[j]);
  int2* range = tl.range + par * tl.ntile;
#pragma unroll
  for (int j = 0; j < kTile1Per; j++) {
    const int r = r0 + j * kBlock + threadIdx.x;
    const int otile = r < min(n, nb0) ? in[j].okey >> tl.tshift : -1;
    // the first and the last record of each tile's source range: its
    // neighbors in the buffer (the next lanes of the wave) belong to another
    // tile, or to none
    int prev = __shfl_up(otile, 1), next = __shfl_down(otile, 1);
    if (lane == 0) prev = (otile >= 0 && r > 0) ? nbk[j] >> tl.tshift : -1;
    if (lane == 63) next = (otile >= 0 && r + 1 < nb0) ? nbk[j] >> tl.tshift : -1;
    if (otile >= 0 && prev != otile) range[otile].x = r;
    if (otile >= 0 && next != otile) range[otile].y = r + 1;
    int key = -1;
    if (r < n) {
      key = tile_finish<kF32>(I, S, a, r, in[j], d[j], gg[j], step, gl);
      tl.nkey[r] = key;
    }
    const bool stay = key >= 0 && (key >> tl.tshift) == otile;
    if (key >= 0 && !stay) tile_queue(tl, tl.imm + par * tl.ntile, tl.nspill + par, key, r);
    wave_run_count(tl.stay + par * tl.ntile, otile, stay);
  }
}

// Each tile's first destination record (one workgroup; every thread takes a
// run of consecutive tiles, eight at a time, with all of a batch's loads
// issued before they are used).
constexpr int kTileScanThreads = 1024;
constexpr int kTileScanBatch = 8;
__global__ __launch_bounds__(kTileScanThreads) void k_tile_scan(const Inst* __restrict__ insts,
                                                                const Glob* __restrict__ gl, Tiles tl) {
  if (live_nrec(gl) < 0) return;
  __shared__ int s_w[kTileScanThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int par = insts[0].ctl->step & 1;
  const int* stay = tl.stay + par * tl.ntile;
  const int* imm = tl.imm + par * tl.ntile;
  const int chunk = (tl.ntile + kTileScanThreads - 1) / kTileScanThreads;
  const int t0 = min(tid * chunk, tl.ntile), t1 = min(t0 + chunk, tl.ntile);
  int v[kTileScanBatch];
  int sum = 0;
  for (int t = t0; t < t1; t += kTileScanBatch) {
#pragma unroll
    for (int j = 0; j < kTileScanBatch; j++) v[j] = t + j < t1 ? stay[t + j] + imm[t + j] : 0;
#pragma unroll
    for (int j = 0; j < kTileScanBatch; j++) sum += v[j];
  }
  const int inc = wave_incl_scan(sum);
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  int run = inc - sum;
  for (int w = 0; w < wv; w++) run += s_w[w];
  for (int t = t0; t < t1; t += kTileScanBatch) {
    if (chunk > kTileScanBatch) {  // the counts of this batch again (the first loop kept only the last batch)
#pragma unroll
      for (int j = 0; j < kTileScanBatch; j++) v[j] = t + j < t1 ? stay[t + j] + imm[t + j] : 0;
    }
#pragma unroll
    for (int j = 0; j < kTileScanBatch; j++) {
      if (t + j < t1) {
        tl.toff[t + j] = run;
        run += v[j];
      }
    }
  }
}

// One record as k_tile2 copies it.
struct TileRec {
  double p[3];  // FP64 position
  float f[3];   // FP32 offsets from the origin of the record's new box
  uint64_t ser;
  int key, ident, lp;
  int rec;  // source record
};

template <bool kF32>
__device__ __forceinline__ void tile_load(const Recs& a, const int* __restrict__ nkey, int r, TileRec& o) {
  o.key = nkey[r];
  o.rec = r;
  o.ident = a.ident[r];
  o.lp = a.lpos[r];
  o.ser = a.serno[r];
  if (kF32) {
    o.f[0] = a.ox[r];
    o.f[1] = a.oy[r];
    o.f[2] = a.oz[r];
  } else {
    o.p[0] = a.px[r];
    o.p[1] = a.py[r];
    o.p[2] = a.pz[r];
  }
}

// The fields a record is written with: its box, list and z sub-bin from its
// key, the position (and the packed FP32 screen copy k_scatter writes).
struct TileOut {
  double x, y, z;  // position (FP64 records: the stored value)
  float4 spk;
  int bx, ll;
};

template <bool kF32>
__device__ __forceinline__ void tile_out(const Inst& I, const TileRec& o, TileOut& v) {
  const Geometry& g = I.g;
  const int zsub = I.zsub;
  const int kk = o.key - I.keybase;
  const int zs = kk & (zsub - 1);
  v.bx = (kk / zsub) % g.nbox;
  v.ll = kk / (zsub * g.nbox);
  if (kF32) {
    double og[3];
    box_origin(g, v.bx, og);
    v.x = og[0] + (double)o.f[0];
    v.y = og[1] + (double)o.f[1];
    v.z = og[2] + (double)o.f[2];
  } else {
    v.x = o.p[0];
    v.y = o.p[1];
    v.z = o.p[2];
  }
  const double lim = I.screen_lim;
  const int far = !(fabs(v.x) <= lim && fabs(v.y) <= lim && fabs(v.z) <= lim);
  v.spk = make_float4((float)v.x, (float)v.y, (float)v.z, __int_as_float(o.ident | (far << 16) | (zs << 17)));
}

// The inverse map (list slot -> record) of the last slots of a list.
__device__ __forceinline__ void tile_inv(const Inst& I, const TileRec& o, const TileOut& v, int dst,
                                         int* __restrict__ inv) {
  if (o.lp >= I.ctl->nl[v.ll] - I.sortcap) inv[I.L.base[v.ll] + o.lp] = dst;
}

// Place a record at dst of the destination buffer, with its packed FP32
// screen copy and, for the last slots of its list, the inverse map.
template <bool kF32>
__device__ __forceinline__ void tile_place(const Inst& I, const TileRec& o, int dst, const Recs& b,
                                           const Binned& bn, int* __restrict__ inv) {
  TileOut v;
  tile_out<kF32>(I, o, v);
  if (kF32) {
    b.ox[dst] = o.f[0];
    b.oy[dst] = o.f[1];
    b.oz[dst] = o.f[2];
  } else {
    b.px[dst] = v.x;
    b.py[dst] = v.y;
    b.pz[dst] = v.z;
  }
  b.ident[dst] = o.ident;
  b.serno[dst] = o.ser;
  b.lpos[dst] = o.lp;
  b.box[dst] = v.bx;
  b.rank[dst] = o.key;
  bn.spk[dst] = v.spk;
  if (I.src) I.src[dst] = o.rec;
  tile_inv(I, o, v, dst, inv);
}

template <bool kF32>
__global__ __launch_bounds__(kBlock) void k_tile2(const Inst* __restrict__ insts, Glob* __restrict__ gl, Recs a,
                                                  Recs b, Binned bn, Tiles tl, int* __restrict__ inv) {
  const int n = live_nrec(gl);
  if (n < 0) return;
  const Inst& I = insts[0];
  __shared__ int s_w[kBlock / 64];
  __shared__ int s_cnt[kTileMaxKeys];   // records per key, then each key's first destination record
  __shared__ int s_more[kTileMaxKeys];  // records per key among the cached ones, then the next free place
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int t = blockIdx.x;
  // the tile's counters of both step parities load with the step counter
  // (no load waits on another)
  const int2 rg0 = tl.range[t], rg1 = tl.range[tl.ntile + t];
  const int in0 = tl.imm[t], in1 = tl.imm[tl.ntile + t];
  const int sp0 = tl.nspill[0], sp1 = tl.nspill[1];
  const int off = tl.toff[t];
  const int par = I.ctl->step & 1;
  const int tk = 1 << tl.tshift, k0 = t << tl.tshift;
  for (int k = tid; k < tk; k += kBlock) s_cnt[k] = 0;
  const int2 rg = par ? rg1 : rg0;
  const int nin = par ? in1 : in0;
  const int nbk = min(nin, kTileBucket);
  const int nsp = nin > kTileBucket ? (par ? sp1 : sp0) : 0;
  const int ns = rg.y - rg.x;
  const int nitems = ns + nbk + nsp;
  __syncthreads();
  // the next step's queue counters (every tile has read this step's above)
  if (tid == 0) {
    tl.imm[(par ^ 1) * tl.ntile + t] = 0;
    tl.stay[(par ^ 1) * tl.ntile + t] = 0;
    tl.range[(par ^ 1) * tl.ntile + t] = make_int2(0, 0);
    if (t == 0) tl.nspill[par ^ 1] = 0;
  }
  // item idx: the tile's source range, then the bucket, then the spill list
  // (entries of other tiles there give -1)
  auto item_rec = [&](int idx) -> int {
    if (idx < ns) return rg.x + idx;
    idx -= ns;
    if (idx < nbk) return tl.bucket[t * kTileBucket + idx];
    const int2 sp = tl.spill[idx - nbk];
    return (sp.y >> tl.tshift) == t ? sp.x : -1;
  };
  TileRec c[kTileCache];
  int rk[kTileCache];
#pragma unroll
  for (int j = 0; j < kTileCache; j++) {
    const int idx = tid + j * kBlock;
    rk[j] = -1;
    const int r = idx < nitems ? item_rec(idx) : -1;
    if (r >= 0) {
      tile_load<kF32>(a, tl.nkey, r, c[j]);
      if (c[j].key >= 0 && (c[j].key >> tl.tshift) == t) rk[j] = atomicAdd(&s_cnt[c[j].key - k0], 1);
    }
  }
  const bool more = nitems > kTileCache * kBlock;  // uniform over the block
  __syncthreads();
  if (more) {  // a crowded tile: the records beyond the cached ones are read again to be placed
    for (int k = tid; k < tk; k += kBlock) s_more[k] = s_cnt[k];
    __syncthreads();
    for (int idx = kTileCache * kBlock + tid; idx < nitems; idx += kBlock) {
      const int r = item_rec(idx);
      if (r < 0) continue;
      const int key = tl.nkey[r];
      if (key >= 0 && (key >> tl.tshift) == t) atomicAdd(&s_cnt[key - k0], 1);
    }
    __syncthreads();
  }
  // exclusive scan of the tile's counts: tk / kBlock consecutive keys per thread
  {
    const int per = tk / kBlock;
    int v[kTileMaxKeys / kBlock];
    int tsum = 0;
#pragma unroll
    for (int q = 0; q < kTileMaxKeys / kBlock; q++) {
      v[q] = q < per ? s_cnt[tid * per + q] : 0;
      tsum += v[q];
    }
    const int inc = wave_incl_scan(tsum);
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    int base = off + inc - tsum, total = 0;
    for (int w = 0; w < kBlock / 64; w++) {
      const int x = s_w[w];
      if (w < wv) base += x;
      total += x;
    }
    const int nk = min(tk, tl.nkeys - k0);
#pragma unroll
    for (int q = 0; q < kTileMaxKeys / kBlock; q++) {
      if (q < per) {
        const int kk = tid * per + q;
        s_cnt[kk] = base;
        if (kk < nk) bn.bstart[k0 + kk] = base;
        base += v[q];
      }
    }
    if (t == tl.ntile - 1 && tid == 0) bn.bstart[tl.nkeys] = off + total;  // records binned
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTileCache; j++)
    if (rk[j] >= 0) tile_place<kF32>(I, c[j], s_cnt[c[j].key - k0] + rk[j], b, bn, inv);
  if (more) {
    for (int k = tid; k < tk; k += kBlock) s_more[k] += s_cnt[k];
    __syncthreads();
    for (int idx = kTileCache * kBlock + tid; idx < nitems; idx += kBlock) {
      const int r = item_rec(idx);
      if (r < 0) continue;
      TileRec o;
      tile_load<kF32>(a, tl.nkey, r, o);
      if (o.key >= 0 && (o.key >> tl.tshift) == t) tile_place<kF32>(I, o, atomicAdd(&s_more[o.key - k0], 1), b, bn, inv);
    }
  }
}

// Full inverse map of the current buffer, for reading the state back.
template <bool kMulti>
__global__ __launch_bounds__(kBlock) void k_rebuild_inv(const Inst* __restrict__ insts, const Glob* __restrict__ gl,
                                                        Recs a, int* __restrict__ inv) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (gl->halt || r >= cur_nrec(insts, gl)) return;
  int i = a.ident[r];
  int in = kMulti ? a.inst[r] : 0;
  if (i <= 0 || i >= kMaxSpecies || in < 0 || in >= gl->ninst) return;
  const Inst& I = insts[in];
  int ll = I.tab->listof[i];
  if (a.sb && a.sb[r]) ll = I.sf.blist[i][bstate(a.sb[r])];
  int lp = a.lpos[r];
  if (lp >= 0 && lp < I.L.base[ll + 1] - I.L.base[ll]) inv[I.L.base[ll] + lp] = r;
}

// ---------------------------------------------------------------------------
// k_pairs: candidate reaction pairs with their reference-order keys

__device__ __forceinline__ void emit_edge(Control* ctl, Edge* edges, int maxedges, uint64_t key, int s1, int s2,
                                          int r, int wrap) {
  int e = atomicAdd(&ctl->nedges, 1);
  if (e < maxedges) {
    Edge ed;
    ed.key = key;
    ed.s1 = s1;
    ed.s2 = s2;
    ed.r = r;
    ed.wrap = wrap;
    edges[e] = ed;
  } else {
    atomicOr(&ctl->err, kErrEdgeOverflow);
  }
}

template <int kSurf>
__device__ __forceinline__ int first_reaction(const DevTables* tab, int i1, int i2, double dist2, uint32_t key0,
                                              uint32_t key1, uint64_t ser1, uint64_t ser2, uint32_t step,
                                              const DevSurf& sf, const Geometry& g, const double* p1,
                                              const double* p2, int wrap) {
  int set = i1 * kMaxSpecies + i2;
  int n = tab->nbi[set];
  for (int j = 0; j < n; j++) {
    int r = tab->bi[set][j];
    const DevRxn& rx = tab->rxn2[r];
    if (!(dist2 <= rx.bindrad2)) continue;
    if (rx.prob != 1.0) {
      smolrng4 rr = smolrng_philox(key0, key1, (uint32_t)ser1, (uint32_t)ser2, step, SMOLRNG_BIRXN | ((uint32_t)j << 8));
      if (!(smolrng_u01(rr.v[0]) < rx.prob)) continue;
    }
    // rxnXsurface (smolreact.c:3242,3289; not for wrapped pairs)
    if (kSurf && !wrap && sf.npnl > 0 &&
        srf_rxnblocked<(kSurf > 1)>(sf, g, tab->difc, p1, p2, i1, i2, ser1, ser2, step, j, key0, key1))
      continue;
    if (!rx.permit_soln) continue;
    return r;
  }
  return -1;
}

// Exact FP64 test of one candidate (the reference's dist2 in its operation
// order, smolreact.c:3238-3240, and wallcalcdist2, smolwall.c:100-116) and
// the edge for the first reaction that accepts it.
template <int kSurf, bool kF32>
__device__ __forceinline__ void exact_pair(const Geometry& g, const DevSurf& sf, const DevTables* tab, Control* ctl,
                                           const Recs& b,
                                           Edge* edges, int maxedges, uint32_t key0, uint32_t key1, uint32_t step,
                                           int t1, const double* p1, int i1, uint64_t ser1, int t2, int i2, int wrap,
                                           uint64_t key) {
  double p2[3];
  load_pos<kF32>(b, g, t2, kF32 ? b.box[t2] : 0, p2);
  double dist2 = 0;
  if (wrap) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      int cc = (wrap >> (2 * d)) & 3;
      double v;
      if (cc == 0) v = p2[d] - p1[d];
      else if (cc == 1) v = p2[d] - p1[d] - g.width[d];
      else v = p2[d] - p1[d] + g.width[d];
      dist2 += v * v;
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; d++) dist2 += (p1[d] - p2[d]) * (p1[d] - p2[d]);
  }
  int r = first_reaction<kSurf>(tab, i1, i2, dist2, key0, key1, ser1, b.serno[t2], step, sf, g, p1, p2, wrap);
  if (r >= 0) emit_edge(ctl, edges, maxedges, key, t1, t2, r, wrap);
}

// Distances of a record's FP32 screen position to the faces of its virtual
// box (x, y) and of its z sub-bin (lower l*, upper u*).  A neighbor column or
// sub-bin lies beyond one of these faces (records are binned by pos2box,
// which clamps only at the grid's ends, where the records beyond still lie
// beyond the face), so when that distance exceeds the screening radius it
// holds no record within the radius and is not scanned.  The radius already
// carries the FP32 margin, which covers the rounding of these distances.
struct Faces {
  float lx, ux, ly, uy, lz, uz;
};

__device__ __forceinline__ Faces box_faces(const Geometry& g, int bx, int by, int bz, int zs, int zsub,
                                           const float4& me) {
  const double xlo = g.bmin[0] + (double)bx * g.bsize[0];
  const double ylo = g.bmin[1] + (double)by * g.bsize[1];
  const double h = g.bsize[2] / (double)zsub;
  const double zlo = g.bmin[2] + (double)bz * g.bsize[2] + (double)zs * h;
  Faces f;
  f.lx = me.x - (float)xlo;
  f.ux = (float)(xlo + g.bsize[0]) - me.x;
  f.ly = me.y - (float)ylo;
  f.uy = (float)(ylo + g.bsize[1]) - me.y;
  f.lz = me.z - (float)zlo;
  f.uz = (float)(zlo + h) - me.z;
  return f;
}

// true when neighbor column (dx, dy) is out of reach
__device__ __forceinline__ bool col_culled(const Faces& f, int dx, int dy, float rf) {
  return (dx < 0 && f.lx > rf) || (dx > 0 && f.ux > rf) || (dy < 0 && f.ly > rf) || (dy > 0 && f.uy > rf);
}

// Pair search, one thread per binned record (box order), molecule m1.
// The stencil is walked as z-runs: in box order the boxes (x', y', z-1..z+1)
// of one neighbor column are consecutive records, so each column is one or
// two contiguous ranges (two when z wraps).  Half stencil (same list): columns
// (-1,-1),(-1,0),(-1,1),(0,-1) with dz -1..1 and column (0,0) with dz -1..0;
// full stencil (two lists): all nine columns, dz -1..1.  dz = 0 in column
// (0,0) is the molecule's own box, bireact's same-box pass.  Neighbor
// columns and z sub-bins whose face lies farther than the screening radius
// from the molecule are skipped (box_faces): most molecules scan only their
// own column's sub-bin and one neighbor sub-bin.  Records that pass the FP32
// screen are handed to k_exact, which runs the reference's FP64 test and
// keys the edge; the screen itself stays small so many waves hide the
// latency of its loads.
// kHalf1: the simulation's only reacting list pair is one list with itself,
// so every molecule that reacts is searched with the half stencil of its own
// list (the column loop unrolls with constant offsets).
template <bool kMulti, int kSurf, bool kHalf1>
__device__ __forceinline__ void screen_global(const Inst* __restrict__ insts, Glob* __restrict__ gl, const Recs& b,
                                              const Binned& bn, int u, int4* __restrict__ cand, int candcap) {
  const Inst& I = insts[kMulti ? b.inst[u] : 0];
  const Geometry& g = I.g;
  const DevTables* tab = I.tab;
  const PairConst& pc = I.pc;
  Control* ctl = I.ctl;
  const int nbox = g.nbox;
  const float4 me = bn.spk[u];
  const int w1 = __float_as_int(me.w);
  const int i1 = w1 & 0xFFFF;
  const uint32_t mask = pc.reactmask[i1];
  if (!mask) return;
  if (kSurf && ((w1 >> 20) & 1)) return;  // surface-bound: no reactions
  const float r2f = ((w1 >> 16) & 1) ? 3.0e38f : pc.r2f[i1];
  const int zsub = I.zsub;
  const int zs1 = (w1 >> 17) & 7;
  const int ll1 = tab->listof[i1];
  const int nl = tab->nlist;
  const int bb = b.box[u];
  const int sz = g.side[2];
  const int bx = bb / (g.side[1] * sz);
  const int by = (bb / sz) % g.side[1];
  const int bz = bb % sz;
  const float rf = sqrtf(r2f) * 1.000001f;
  const Faces fc = box_faces(g, bx, by, bz, zs1, zsub, me);
  auto column = [&](const int c, const int dx, const int dy, const int ll2, const bool half) {
    if (col_culled(fc, dx, dy, rf)) return;
    int nx = bx + dx, ny = by + dy, cw = 0;
    float shx = 0.f, shy = 0.f;
    if (nx < 0) {
      if (!g.wrapflag[0]) return;
      nx = g.side[0] - 1;
      cw |= 1;
      shx = -(float)g.width[0];
    } else if (nx >= g.side[0]) {
      if (!g.wrapflag[0]) return;
      nx = 0;
      cw |= 2;
      shx = (float)g.width[0];
    }
    if (ny < 0) {
      if (!g.wrapflag[1]) return;
      ny = g.side[1] - 1;
      cw |= 1 << 2;
      shy = -(float)g.width[1];
    } else if (ny >= g.side[1]) {
      if (!g.wrapflag[1]) return;
      ny = 0;
      cw |= 2 << 2;
      shy = (float)g.width[1];
    }
    const int colb = I.keybase + (ll2 * nbox + (nx * g.side[1] + ny) * sz) * zsub;
    // z segments in sub-bins: unwrapped run [zlo, zhi] plus wrapped single
    // sub-bins; column (0,0) of the half stencil stops at its own box
    const int nzs = sz * zsub;
    const int gz1 = bz * zsub + zs1;
    // sub-bins below and above only when their face is within reach
    const int zlo = fc.lz > rf ? gz1 : gz1 - 1;
    const int zup = fc.uz > rf ? gz1 : gz1 + 1;
    const int zhi = (half && c == 4) ? min(zup, (bz + 1) * zsub - 1) : zup;
    for (int seg = 0; seg < 3; seg++) {
      int za, zb2, zw = 0, dzfix = -2;
      float shz = 0.f;
      if (seg == 0) {  // wrapped box below
        if (zlo >= 0) continue;
        if (!g.wrapflag[2]) continue;
        za = zb2 = nzs - 1;
        zw = 1 << 4;
        shz = -(float)g.width[2];
        dzfix = -1;
      } else if (seg == 1) {  // unwrapped run
        za = zlo < 0 ? 0 : zlo;
        zb2 = zhi >= nzs ? nzs - 1 : zhi;
        if (za > zb2) continue;
      } else {  // wrapped box above
        if (zhi < nzs) continue;
        if (!g.wrapflag[2]) continue;
        za = zb2 = 0;
        zw = 2 << 4;
        shz = (float)g.width[2];
        dzfix = 1;
      }
      const int k0 = bn.bstart[colb + za];
      const int k1 = bn.bstart[colb + zb2 + 1];
      const int wrap = cw | zw;
      const float ox = me.x - shx, oy = me.y - shy, oz = me.z - shz;
      for (int kb = k0; kb < k1; kb += 32) {
        // screening: bit j of hits <=> record kb+j is a candidate partner
        const int kn = min(k1 - kb, 32);
        uint32_t hits = 0;
#pragma unroll 4
        for (int j = 0; j < kn; j++) {
          const float4 o = bn.spk[kb + j];
          const float dxf = o.x - ox, dyf = o.y - oy, dzf = o.z - oz;
          const float d2 = dxf * dxf + dyf * dyf + dzf * dzf;
          const int w2 = __float_as_int(o.w);
          bool hit = ((d2 <= r2f) | ((w2 >> 16) & 1)) & ((mask >> (w2 & 31)) & 1u);
          if (kSurf) hit = hit & !((w2 >> 20) & 1);
          hits |= (uint32_t)hit << j;
        }
        if (u >= kb && u < kb + 32) hits &= ~(1u << (u - kb));
        // candidates go to k_exact: record pair and the stencil position
        // (column offsets, fixed dz of a wrapped sub-bin, wrap code, list)
        const int code = (dx + 1) | ((dy + 1) << 2) | ((dzfix + 2) << 4) | (wrap << 6) | (ll2 << 12) |
                         ((int)half << 15);
        while (hits) {
          const int k = kb + __ffs(hits) - 1;
          hits &= hits - 1;
          const int ci = atomicAdd(&gl->ncand, 1);
          if (ci < candcap) cand[ci] = make_int4(u, k, code, 0);
          else atomicOr(&ctl->err, kErrEdgeOverflow);
        }
      }
    }
  };
  if (kHalf1) {  // the half stencil in the reference's column order
    column(0, -1, -1, ll1, true);
    column(1, -1, 0, ll1, true);
    column(2, -1, 1, ll1, true);
    column(3, 0, -1, ll1, true);
    column(4, 0, 0, ll1, true);
  } else {
    for (int ll2 = ll1; ll2 < nl; ll2++) {
      if (!tab->rxnmol2[ll1 * kMaxList + ll2]) continue;
      const bool half = (ll1 == ll2);
      const int ncol = half ? 5 : 9;
      for (int c = 0; c < ncol; c++)
        column(c, half ? (c < 3 ? -1 : 0) : c / 3 - 1, half ? (c < 3 ? c - 1 : c - 4) : c % 3 - 1, ll2, half);
    }
  }
}

template <bool kMulti, int kSurf, bool kHalf1>
__global__ __launch_bounds__(kBlock) void k_pairs(const Inst* __restrict__ insts, Glob* __restrict__ gl, Recs b,
                                                  Binned bn, int nscan, int4* __restrict__ cand, int candcap) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (gl->halt || u >= bn.bstart[nscan - 1]) return;  // records binned
  screen_global<kMulti, kSurf, kHalf1>(insts, gl, b, bn, u, cand, candcap);
}

// Exact test of the screened candidates (one thread per candidate, grid-
// stride): stencil index and edge key of molecule 1's loop position, then the
// reference's FP64 distance test and reaction choice (exact_pair).
template <int kSurf, bool kMulti, bool kF32>
__global__ __launch_bounds__(kBlock) void k_exact(const Inst* __restrict__ insts, const Glob* __restrict__ gl, Recs b,
                                                  Binned bn, const int4* __restrict__ cand, int candcap) {
  if (gl->halt) return;
  const int n = min(gl->ncand, candcap);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int4 cd = cand[c];
    const int u = cd.x, k = cd.y, code = cd.z;
    const int dx = (code & 3) - 1, dy = ((code >> 2) & 3) - 1, dzfix = ((code >> 4) & 3) - 2;
    const int wrap = (code >> 6) & 63, ll2 = (code >> 12) & 7;
    const bool half = (code >> 15) & 1;
    const Inst& I = insts[kMulti ? b.inst[u] : 0];
    const Geometry& g = I.g;
    const DevTables* tab = I.tab;
    const int sz = g.side[2];
    const int bb = b.box[u];
    const int dz = dzfix > -2 ? dzfix : (b.box[k] % sz) - bb % sz;
    const int nbi = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    // molecule 1's live position orders the outer loop; the partners come in
    // box-list order (= live order unless this step's first molsort changed
    // the box list: surfaces)
    const int lp1 = b.lpos[u];
    const int o1 = kSurf ? b.ov[u] : lp1, lp2 = kSurf ? b.ov[k] : b.lpos[k];
    if (nbi == 13 && half && lp2 >= o1) continue;
    const int i1 = __float_as_int(bn.spk[u].w) & 0xFFFF;
    const int i2 = __float_as_int(bn.spk[k].w) & 0xFFFF;
    const int ll1 = tab->listof[i1];
    const uint64_t key1base = ((uint64_t)tab->pairindex[ll1 * kMaxList + ll2] << 59) | ((uint64_t)lp1 << 32);
    const uint64_t key = nbi == 13 ? (key1base | (uint64_t)lp2)
                                   : ((1ull << 63) | key1base | ((uint64_t)nbi << 27) | (uint64_t)lp2);
    double p1[3];
    load_pos<kF32>(b, g, u, bb, p1);
    exact_pair<kSurf, kF32>(g, I.sf, tab, I.ctl, b, I.edges, I.maxedges, I.key0, I.key1, I.ctl->step, u, p1, i1,
                            b.serno[u], k, i2, nbi == 13 ? 0 : wrap, key);
  }
}

// List slot of a first-order reactant at the time of unireact: its live
// position after the step's first molsort (the record of the source buffer).
__device__ __forceinline__ int uni_slot(const Inst& I, const Recs& RA, int rec) {
  return I.L.base[I.tab->listof[RA.ident[rec]]] + RA.lpos[rec];
}

// Geo_SphereReflectSphere (source/libSteve/Geometry.c:1264-1293) in three
// dimensions: two molecules move on straight lines a0->a1 and b0->b1; p is
// the fraction of the step at which their separation first equals the
// binding radius (radius2 = radius squared); each molecule's remaining motion
// is reflected off the plane normal to their separation at that moment, the
// reflected end points in a1p and b1p.  Returns p.
__device__ __forceinline__ double sphere_reflect_sphere(const double* a0, const double* a1, const double* b0,
                                                        const double* b1, double radius2, double* a1p, double* b1p) {
  double qa = 0, qb = 0, qc = 0, deltap[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double rel = b1[d] - a1[d] - b0[d] + a0[d];  // change of the separation over the step
    const double sep0 = b0[d] - a0[d];                  // separation at the start
    qa += rel * rel;
    qb += 2 * sep0 * rel;
    qc += sep0 * sep0;
  }
  qc -= radius2;
  const double p = (-qb - sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
  double f1 = 0, f2 = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    deltap[d] = (1 - p) * (b0[d] - a0[d]) + p * (b1[d] - a1[d]);
    f1 += (a1[d] - a0[d]) * deltap[d];
    f2 += (b1[d] - b0[d]) * deltap[d];
  }
  const double k = 2 * (1 - p) / radius2;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a1p[d] = f1 >= 0 ? a1[d] - k * f1 * deltap[d] : 2 * (1 - p) * a0[d] + (2 * p - 1) * a1[d] + k * f1 * deltap[d];
    b1p[d] = f2 <= 0 ? b1[d] - k * f2 * deltap[d] : 2 * (1 - p) * b0[d] + (2 * p - 1) * b1[d] + k * f2 * deltap[d];
  }
  return p;
}

// ---------------------------------------------------------------------------
// k_resolve helpers (one workgroup)

// Bitonic sort of (key, val) pairs, n a power of two, in LDS or global memory.
template <typename KP, typename VP>
__device__ void wg_bitonic(KP key, VP val, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t x = key[i], c = key[ixj];
          bool up = ((i & k) == 0);
          if ((x > c) == up) {
            key[i] = c;
            key[ixj] = x;
            int tv = val[i];
            val[i] = val[ixj];
            val[ixj] = tv;
          }
        }
      }
      __syncthreads();
    }
  }
}

__host__ __device__ inline int pow2ceil(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Sort n (key,val) pairs held in gk/gv (global, capacity >= pow2ceil(n)).
__device__ void wg_sort(uint64_t* gk, int* gv, int n, uint64_t* lk, int* lv) {
  if (n <= 1) return;
  int np = pow2ceil(n);
  if (np <= kLdsSort) {
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
      lk[i] = i < n ? gk[i] : kNoKey;
      lv[i] = i < n ? gv[i] : 0;
    }
    __syncthreads();
    wg_bitonic(lk, lv, np);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      gk[i] = lk[i];
      gv[i] = lv[i];
    }
    __syncthreads();
  } else {
    for (int i = n + threadIdx.x; i < np; i += blockDim.x) {
      gk[i] = kNoKey;
      gv[i] = 0;
    }
    __syncthreads();
    wg_bitonic(gk, gv, np);
  }
}

// Sort n <= kRankSort (key,val) pairs in gk/gv in place by rank counting in
// LDS: each key's destination is the number of keys before it (ties by
// index).  Two barriers instead of the log^2 stages of a bitonic network.
constexpr int kRankSort = 512;
// Sort n <= blockDim int keys of k_in (LDS) into k_out (LDS), ascending.
// Keys are spread over kBuckets buckets by their place in [min, max]; the
// bucket counts (LDS atomics) are scanned, each key is placed in its bucket,
// and its rank is the bucket's start plus the keys of the bucket below it
// (equal keys by their place in the bucket).  Five barriers and, for keys
// spread like list slots, about n / kBuckets comparisons per key.
// scratch: LDS, kBuckets + 1 + n + 2 * (blockDim / 64) ints.
constexpr int kBuckets = 256;
__device__ void wg_bucket_sort(const int* k_in, int n, int* k_out, int* scratch) {
  int* cnt = scratch;                 // kBuckets + 1
  int* tmp = scratch + kBuckets + 1;  // n
  int* red = tmp + n;                 // min, max of each wave
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
  int key = 0, mn = INT_MAX, mx = INT_MIN;
  if (tid < n) {
    key = k_in[tid];
    mn = key;
    mx = key;
  }
  for (int i = tid; i <= kBuckets; i += blockDim.x) cnt[i] = 0;
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, __shfl_xor(mn, o));
    mx = max(mx, __shfl_xor(mx, o));
  }
  if (lane == 0) {
    red[2 * w] = mn;
    red[2 * w + 1] = mx;
  }
  __syncthreads();
  mn = red[0];
  mx = red[1];
  for (int v = 1; v < nw; v++) {
    mn = min(mn, red[2 * v]);
    mx = max(mx, red[2 * v + 1]);
  }
  const long long span = (long long)mx - mn + 1;
  int b = 0, pos = 0;
  if (tid < n) {
    b = (int)(((long long)key - mn) * kBuckets / span);
    pos = atomicAdd(&cnt[b], 1);
  }
  __syncthreads();
  if (w == 0) {  // exclusive scan of the counts, kBuckets / 64 per lane
    constexpr int per = kBuckets / 64;
    int c[per], t = 0;
#pragma unroll
    for (int u = 0; u < per; u++) {
      c[u] = cnt[per * lane + u];
      t += c[u];
    }
    int inc = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    int run = inc - t;
#pragma unroll
    for (int u = 0; u < per; u++) {
      cnt[per * lane + u] = run;
      run += c[u];
    }
    if (lane == 63) cnt[kBuckets] = run;
  }
  __syncthreads();
  int st = 0;
  if (tid < n) {
    st = cnt[b];
    tmp[st + pos] = key;
  }
  __syncthreads();
  if (tid < n) {
    const int en = cnt[b + 1], me = st + pos;
    int r = st;
    for (int j = st; j < en; j++) {
      const int x = tmp[j];
      r += (x < key) | ((x == key) & (j < me));
    }
    k_out[r] = key;
  }
  __syncthreads();
}

// Rank of key k (element i) among lk[j0, j1): keys below it, and equal keys
// at lower indices.  The LDS loads go out eight at a time, so the loop waits
// once per eight keys instead of once per key.
template <typename K>
__device__ __forceinline__ int rank_in(const K* lk, int j0, int j1, int i, K k) {
  int r = 0, j = j0;
  for (; j + 8 <= j1; j += 8) {
    K x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = lk[j + u];
#pragma unroll
    for (int u = 0; u < 8; u++) r += (x[u] < k) | ((x[u] == k) & (j + u < i));
  }
  for (; j < j1; j++) {
    const K x = lk[j];
    r += (x < k) | ((x == k) & (j < i));
  }
  return r;
}

__device__ void wg_rank_sort(uint64_t* gk, int* gv, int n, uint64_t* lk, int* lv, int* ok32, int* ov) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    lk[i] = gk[i];
    lv[i] = gv[i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint64_t k = lk[i];
    const int r = rank_in<uint64_t>(lk, 0, n, i, k);
    gk[r] = k;
    gv[r] = lv[i];
    if (ok32) ok32[r] = (int)k;
    if (ov) ov[r] = lv[i];
  }
  __syncthreads();
}

// Rank sort, in LDS, of two independent groups [0, n0) and [n0, n0+n1) of
// (key, val) pairs already in lk/lv: each group is ordered by key (ties by
// index) and its sorted values and low key words go to ov / ok32 (either may
// be null), group 1 after group 0.  Up to 64 threads share the comparisons of
// one element (a power of two, so their lanes are aligned within a wave).
template <typename K>
__device__ void wg_rank_sort2(const K* lk, const int* lv, int n0, int n1, int* ov, int* ok32) {
  const int n = n0 + n1;
  int tpe = 1;
  while (tpe < 64 && n * tpe * 2 <= (int)blockDim.x) tpe <<= 1;
  const int total = n * tpe;
  for (int base = 0; base < total; base += blockDim.x) {  // same trip count for every thread
    const int t = base + threadIdx.x;
    const int i = t / tpe, q = t & (tpe - 1);
    int r = 0;
    K k = 0;
    if (i < n) {
      k = lk[i];
      // this thread's contiguous share of the element's group
      const int lo = i < n0 ? 0 : n0, hi = i < n0 ? n0 : n;
      const int c = (hi - lo + tpe - 1) / tpe;
      const int j0 = lo + q * c;
      r = rank_in<K>(lk, j0, j0 + c < hi ? j0 + c : hi, i, k);
    }
    for (int o = tpe >> 1; o > 0; o >>= 1) r += __shfl_down(r, o, tpe);
    if (i < n && q == 0) {
      const int dst = (i < n0 ? 0 : n0) + r;
      if (ov) ov[dst] = lv[i];
      if (ok32) ok32[dst] = (int)k;
    }
  }
  __syncthreads();
}

// Sort n (key,val) pairs of gk/gv in place; when ok32/ov are given (LDS,
// capacity >= n) the sorted low key words / values are also left there.
__device__ void wg_sort_any(uint64_t* gk, int* gv, int n, uint64_t* lk, int* lv, int* ok32 = nullptr,
                            int* ov = nullptr) {
  if (n <= kRankSort) {
    if (n == 1 && (ok32 || ov)) {
      if (threadIdx.x == 0) {
        if (ok32) ok32[0] = (int)gk[0];
        if (ov) ov[0] = gv[0];
      }
      __syncthreads();
      return;
    }
    if (n >= 1) wg_rank_sort(gk, gv, n, lk, lv, ok32, ov);
    return;
  }
  wg_sort(gk, gv, n, lk, lv);
  if (ok32 || ov) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      if (ok32) ok32[i] = (int)gk[i];
      if (ov) ov[i] = gv[i];
    }
    __syncthreads();
  }
}

// exclusive scan of one value per thread (wave scans by shuffles, then the
// wave totals through LDS); returns the prefix, total in *tot.  s holds at
// least blockDim.x/64 ints.
__device__ int wg_scan(int v, int* s, int* tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) s[w] = inc;
  __syncthreads();
  int off = 0, t = 0;
  for (int i = 0; i < nw; i++) {
    int x = s[i];
    off += (i < w) ? x : 0;
    t += x;
  }
  *tot = t;
  __syncthreads();
  return off + inc - v;
}

constexpr int kLdsEvents = 1024;  // sorted event indices kept in LDS up to this count
constexpr int kLdsKills = 2048;   // sorted kill slots kept in LDS up to this count
constexpr int kLdsMatch = 512;    // edges matched in LDS (more: the same rounds on global arrays)
constexpr int kMatchHash = 2048;  // LDS slots of their molecules (at least twice the molecules)
static_assert(kMatchHash <= kLdsSort && kMatchHash <= kLdsKills && kMatchHash >= 4 * kLdsMatch,
              "LDS matching reuses the sort and kill arrays");
static_assert(kRankSort <= kResolveThreads && kRankSort <= kLdsKills && kBuckets % 64 == 0 &&
                  kBuckets + 1 + kRankSort + 2 * (kResolveThreads / 64) <= 2 * kLdsSort,
              "the kill sort takes one key per thread and its scratch from the sort space");
constexpr int kMaxAds = 512;      // molecules leaving their lists in the sort that ends a step
constexpr int kProdKilled = -1;   // Prod::sb of a product its reborn surface check absorbed

// Sorted kill slots, read from LDS when they fit there, else from global.
struct KillView {
  const int* lds;
  const uint64_t* glb;
  __device__ __forceinline__ int operator[](int i) const { return lds ? lds[i] : (int)glb[i]; }
};


// First molsort of a step with surfaces (smolmolec.c:2943-3023, called after
// checksurfaces and checksurfacebound, smolsim.cpp:2618-2636): applies the
// list changes k_diffuse queued, in the reference's order, to the live
// positions (lpos) and the inverse map, before the pair search keys them.
// Per list in order: the removed slots (molecules that moved out or died)
// are filled from the top of the list as molsort's swap-from-tail scan does;
// molecules moved in by earlier lists sit after the list and are rotated by
// one per removal (the scan's second swap); removed molecules are met in the
// scan's order and a moving one is appended to its new list then.  Last, the
// desorbed molecules (the reference's resurrected ones, one serial number
// each taken from the counter) are appended in the dead list's order:
// reverse creation order rotated by the number of absorbed molecules (every
// kill of the scan rotates the resurrected entries by one; the desorbed
// originals' kills rotate them back to where they were).  One workgroup per
// simulation; the placement itself is sequential, in LDS, on few events.
constexpr int kRelistMax = 512;
constexpr int kAppendOv = 1 << 26;  // order values of box-list appends: above every live position
__global__ __launch_bounds__(kResolveThreads) void k_relist(const Inst* __restrict__ insts, Glob* gl, Recs RA,
                                                            Recs RB, int* __restrict__ inv, int reccap, Binned bn) {
  const Inst& I = insts[blockIdx.x];
  if (gl->halt) return;
  Control* ctl = I.ctl;
  const int M0 = ctl->nmove;
  if (M0 == 0) return;  // uniform over the workgroup
  __shared__ int s_key[kRelistMax], s_val[kRelistMax], s_slot[kRelistMax], s_ev[kRelistMax], s_top[kRelistMax];
  __shared__ int s_app[kRelistMax], s_appl[kRelistMax];
  __shared__ int s_pref[2 * kRelistMax], s_plist[2 * kRelistMax], s_ppos[2 * kRelistMax];
  __shared__ int s_lo[kMaxList], s_cnt[kMaxList], s_nl[kMaxList], s_np, s_nop;
  // box-list operations in the scan's order: key = ((list*nbox + box) << 11) | sequence, value = event | kind<<30
  __shared__ uint64_t s_opk[2 * kRelistMax];
  __shared__ int s_opv[2 * kRelistMax], s_ops[2 * kRelistMax], s_ord[2 * kRelistMax];
  const int tid = threadIdx.x, NT = blockDim.x;
  const int M = M0 < I.maxevents ? M0 : I.maxevents;
  if (M > kRelistMax) {
    if (tid == 0) {
      atomicOr(&ctl->err, kErrSortOverflow);
      atomicOr(&gl->halt, 1);
    }
    return;
  }
  const int nlist = I.L.nlist;
  for (int e = tid; e < M; e += NT) {
    s_key[e] = I.moves[e].slot;
    s_val[e] = e;
  }
  __syncthreads();
  wg_rank_sort2<int>(s_key, s_val, M, 0, s_ev, s_slot);  // events by old slot
  if (tid < nlist) {  // each list's removed slots: [s_lo, s_lo + s_cnt) of the sorted events
    const int lo = I.L.base[tid], hi = I.L.base[tid + 1];
    int l = 0, h = M;
    while (l < h) {
      const int mid = (l + h) >> 1;
      if (s_slot[mid] < lo) l = mid + 1;
      else h = mid;
    }
    const int k0 = l;
    h = M;
    while (l < h) {
      const int mid = (l + h) >> 1;
      if (s_slot[mid] < hi) l = mid + 1;
      else h = mid;
    }
    s_lo[tid] = k0;
    s_cnt[tid] = l - k0;
    s_nl[tid] = ctl->nl[tid];
  }
  __syncthreads();
  // records at the top r slots of each list (r = its removals): the kept
  // ones fill the holes
  for (int ll = 0; ll < nlist; ll++) {
    const int lo = s_lo[ll], r = s_cnt[ll], b = s_nl[ll] - r;
    for (int k = tid; k < r; k += NT) s_top[lo + k] = inv[I.L.base[ll] + b + k];
  }
  __syncthreads();
  if (tid == 0) {
    int np = 0, napp = 0, ndes = 0, nop = 0;
    const uint64_t nbx = (uint64_t)I.g.nbox;
    auto boxop = [&](int list, int box, int e, int add) {  // boxremovemol / boxaddmol of the scan
      s_opk[nop] = (((uint64_t)list * nbx + (uint64_t)box) << 11) | (uint64_t)nop;
      s_opv[nop] = e | (add << 30);
      nop++;
    };
    auto ref = [&](int e) { return I.moves[e].dst >= 0 ? I.moves[e].dst : -(I.moves[e].rec + 2); };
    auto place = [&](int rf, int list, int pos) {
      s_pref[np] = rf;
      s_plist[np] = list;
      s_ppos[np] = pos;
      np++;
    };
    for (int ll = 0; ll < nlist; ll++) {
      const int lo = s_lo[ll], r = s_cnt[ll], n0 = s_nl[ll], b = n0 - r, base = I.L.base[ll];
      auto meet = [&](int e) {  // a removed molecule met by the scan
        const MoveEv mv = I.moves[e];
        boxop(ll, mv.obox, e, 0);
        if (mv.kind != kMoveList) return;  // a desorbed molecule: its slot dies here
        boxop(mv.list, mv.nbox, e, 1);
        if (mv.list > ll) {
          s_app[napp] = e;
          s_appl[napp] = mv.list;
          napp++;
        } else {
          place(ref(e), mv.list, s_nl[mv.list]++);
        }
      };
      int nlow = 0;
      while (nlow < r && s_slot[lo + nlow] - base < b) nlow++;
      int tp = n0, rtop = lo + r - 1;
      for (int j = 0; j < nlow; j++) {
        meet(s_ev[lo + j]);
        for (;;) {
          if (tp <= b) break;  // only on inconsistent input; the hole stays unfilled
          tp--;
          if (rtop >= lo + nlow && s_slot[rtop] - base == tp) {
            meet(s_ev[rtop]);
            rtop--;
            continue;
          }
          place(s_top[lo + (tp - b)], ll, s_slot[lo + j] - base);
          break;
        }
      }
      if (tp > b) {  // the rest of the top slots are all removed: met from b, then downward
        if (nlow < r) meet(s_ev[lo + nlow]);
        for (int q = tp - 1; q > b; q--)
          if (nlow + (q - b) < r) meet(s_ev[lo + nlow + (q - b)]);
      }
      // molecules appended to this list by earlier lists, rotated by r
      int a = 0;
      for (int k = 0; k < napp; k++) a += s_appl[k] == ll;
      if (a > 0) {
        int idx = 0;
        for (int k = 0; k < napp; k++) {
          if (s_appl[k] != ll) continue;
          const int fin = (idx + r) % a;  // the item at append index idx ends at b + fin
          place(ref(s_app[k]), ll, b + fin);
          idx++;
        }
      }
      s_nl[ll] = b + a;
    }
    // desorbed molecules: the resurrected region of the dead list holds them
    // in reverse creation order (creation order = checksurfacebound's, i.e.
    // by old slot); every kill of the scan (desorbed originals and absorbed
    // molecules) moves the region's first entry to its end, so they are
    // appended rotated by the absorbed count
    int nkill = 0;
    for (int k = 0; k < M; k++) {
      const int kd = I.moves[s_ev[k]].kind;
      ndes += kd == kMoveDesorb;
      nkill += kd == kMoveKill;
    }
    if (ndes > 0) {
      int j = 0, first = nkill % ndes;  // index in reverse creation order of the first one appended
      for (int pass = 0; pass < 2; pass++)
        for (int k = M - 1; k >= 0; k--) {
          const int e = s_ev[k];
          if (I.moves[e].kind != kMoveDesorb) continue;
          const int idx = j++ - pass * ndes;  // position in reverse creation order, once per pass
          if (pass == 0 ? idx < first : idx >= first) continue;
          place(ref(e), I.moves[e].list, s_nl[I.moves[e].list]++);
          boxop(I.moves[e].list, I.moves[e].nbox, e, 1);
        }
    }
    s_np = np;
    s_nop = nop;
    for (int ll = 0; ll < nlist; ll++) ctl->nl[ll] = s_nl[ll];
    ctl->serno_next += (uint64_t)ndes;
    ctl->nmove = 0;
  }
  __syncthreads();
  for (int k = tid; k < s_np; k += NT) {
    const int rf = s_pref[k], list = s_plist[k], pos = s_ppos[k];
    const bool ok = (rf >= 0 ? rf < reccap : (rf <= -2 && -rf - 2 < reccap)) && list >= 0 && list < nlist &&
                    pos >= 0 && pos < I.L.base[list + 1] - I.L.base[list];
    if (!ok) {  // inconsistent list state: stop rather than write out of range
      atomicOr(&ctl->err, kErrCapacity);
      atomicOr(&gl->halt, 1);
      continue;
    }
    if (rf >= 0) RB.lpos[rf] = pos;
    else RA.lpos[-rf - 2] = pos;
    inv[I.L.base[list] + pos] = rf;
  }
  // box-list order after the scan: bireact walks box lists, which reassign
  // filled in live order (order value = old live position) and which the scan
  // then changed: a removal moves the box list's last molecule into the
  // removed one's place (boxremovemol, smolboxes.c:167-175; it takes that
  // order value), an addition goes last (boxaddmol).  Operations grouped by
  // (list, box) in the scan's order; one thread per group.
  const int nop = s_nop;
  for (int k = tid; k < nop; k += NT) s_ops[k] = k;
  __syncthreads();
  wg_rank_sort2<uint64_t>(s_opk, s_ops, nop, 0, s_ord, nullptr);
  const int* ord = s_ord;  // op indices by (list, box, sequence)
  for (int k = tid; k < nop; k += NT) {
    const uint64_t grp = s_opk[ord[k]] >> 11;
    if (k > 0 && (s_opk[ord[k - 1]] >> 11) == grp) continue;  // not the first op of its group
    const int L = (int)(grp / (uint64_t)I.g.nbox), X = (int)(grp % (uint64_t)I.g.nbox);
    int kend = k;
    while (kend < nop && (s_opk[ord[kend]] >> 11) == grp) kend++;
    const int key0 = I.keybase + (L * I.g.nbox + X) * I.zsub;
    const int r0 = bn.bstart[key0], r1 = bn.bstart[key0 + I.zsub];
    // first-order reactants of this step in this box (unbinned members), and
    // unbinned molecules added to it
    constexpr int kLoc = 32;
    int lov[kLoc], nloc = 0;
    const int nu = min(ctl->nuni, I.maxevents);
    for (int u = 0; u < nu; u++) {
      const UniEvent ev = I.uniev[u];
      if (ev.slot < I.L.base[L] || ev.slot >= I.L.base[L + 1] || ev.box != X) continue;
      if (nloc < kLoc) lov[nloc++] = ev.slot - I.L.base[L];
      else atomicOr(&ctl->err, kErrCapacity);
    }
    for (int q = k; q < kend; q++) {
      const int v = s_opv[ord[q]], e = v & 0x3FFFFFFF, add = v >> 30;
      const int seq = (int)(s_opk[ord[q]] & 2047);
      if (add) {
        const int dst = I.moves[e].dst;
        if (dst >= 0) RB.ov[dst] = kAppendOv + seq;
        else if (nloc < kLoc) lov[nloc++] = kAppendOv + seq;
        else atomicOr(&ctl->err, kErrCapacity);
        continue;
      }
      // removal of e: the member with the largest order value takes its value
      const int ovx = I.moves[e].ov;
      int best = -1, bkind = 0, bidx = -1;
      for (int rr = r0; rr < r1; rr++) {
        const int o = RB.ov[rr];
        if (o > best) {
          best = o;
          bkind = 1;
          bidx = rr;
        }
      }
      for (int t = 0; t < nloc; t++)
        if (lov[t] > best) {
          best = lov[t];
          bkind = 2;
          bidx = t;
        }
      for (int q2 = q + 1; q2 < kend; q2++) {  // molecules still to leave this box list
        const int v2 = s_opv[ord[q2]];
        if (v2 >> 30) continue;
        const int o = I.moves[v2 & 0x3FFFFFFF].ov;
        if (o > best) {
          best = o;
          bkind = 3;
          bidx = v2 & 0x3FFFFFFF;
        }
      }
      if (best > ovx) {
        if (bkind == 1) RB.ov[bidx] = ovx;
        else if (bkind == 2) lov[bidx] = ovx;
        else I.moves[bidx].ov = ovx;
      }
    }
  }
}

// The list sort that ends every step (simdocommands, smolsim.cpp:2571-2581,
// a molsort of smolmolec.c:2943-3023) for products that their reborn surface
// check adsorbed (molchangeident, smolmolec.c:2902-2906, put them in the list
// of their bound state) or absorbed (molkill).  Those products sit among the
// products just appended to the end of their solution lists, and nothing
// else changes list in this sort.  Per list in order, the scan from the
// lowest such slot meets each one, appends it to its bound list (an absorbed
// one goes nowhere) and fills its hole from the top of
// the list (swap from tail); holes are met in ascending order, each one
// followed by the removed molecules the swaps bring into it; when the
// remaining top slots are all removed they are met from the lowest upward
// and then downward from the top.  One thread runs the scan on the few
// adsorbed products; the writes are spread over the workgroup.
// ws_key, ws_val: LDS scratch of at least 2 * kMaxAds ints each; ws_aux: 6 * kMaxAds.
constexpr int kNoSlot = -1;
__device__ void end_of_step_sort(const Inst& I, Control* ctl, Glob* gl, const Recs& B, int* __restrict__ inv,
                                 int na, const int* s_aslot, const int* s_arec, const int* s_alist, const int* s_nf,
                                 int* s_pl, int* ws_key, int* ws_val, int* ws_aux) {
  const int tid = threadIdx.x, NT = blockDim.x;
  if (na == 0) return;  // uniform: s_nads is read after a barrier
  if (na > kMaxAds) {
    if (tid == 0) {
      atomicOr(&ctl->err, kErrSortOverflow);
      atomicOr(&gl->halt, 1);
    }
    return;
  }
  int* s_sorted = ws_key;            // entry indices by slot
  int* s_sslot = ws_key + kMaxAds;   // their slots
  int* s_top = ws_val;               // records at the top slots of each list
  int* s_tent = ws_val + kMaxAds;    // entry of a top slot that is removed, else -1
  int* s_prec = ws_aux;              // placements: record, list, position
  int* s_plist = ws_aux + 2 * kMaxAds;
  int* s_ppos = ws_aux + 4 * kMaxAds;
  __shared__ int s_key[kMaxAds], s_ord[kMaxAds], s_len[kMaxList], s_lo[kMaxList], s_cnt[kMaxList], s_np;
  for (int a = tid; a < na; a += NT) {
    s_key[a] = s_aslot[a];
    s_ord[a] = a;
  }
  __syncthreads();
  wg_rank_sort2<int>(s_key, s_ord, na, 0, s_sorted, s_sslot);
  const int nlist = I.L.nlist;
  if (tid < nlist) {
    const int lo = I.L.base[tid], hi = I.L.base[tid + 1];
    int l = 0, h = na;
    while (l < h) {
      const int mid = (l + h) >> 1;
      if (s_sslot[mid] < lo) l = mid + 1;
      else h = mid;
    }
    const int k0 = l;
    h = na;
    while (l < h) {
      const int mid = (l + h) >> 1;
      if (s_sslot[mid] < hi) l = mid + 1;
      else h = mid;
    }
    s_lo[tid] = k0;
    s_cnt[tid] = l - k0;
    s_len[tid] = s_nf[tid] + s_pl[tid];
  }
  __syncthreads();
  for (int ll = 0; ll < nlist; ll++) {  // top r slots of each list and which of them leave
    const int lo = s_lo[ll], r = s_cnt[ll], b = s_len[ll] - r, base = I.L.base[ll];
    for (int q = tid; q < r; q += NT) {
      s_top[lo + q] = inv[base + b + q];
      s_tent[lo + q] = kNoSlot;
    }
  }
  __syncthreads();
  for (int ll = 0; ll < nlist; ll++) {
    const int lo = s_lo[ll], r = s_cnt[ll], b = s_len[ll] - r, base = I.L.base[ll];
    for (int k = tid; k < r; k += NT) {
      const int t = s_sslot[lo + k] - base;
      if (t >= b) s_tent[lo + (t - b)] = s_sorted[lo + k];
    }
  }
  __syncthreads();
  if (tid == 0) {
    int np = 0;
    auto place = [&](int rec, int list, int pos) {
      s_prec[np] = rec;
      s_plist[np] = list;
      s_ppos[np] = pos;
      np++;
    };
    auto meet = [&](int e) {  // append to the list of its bound state, or die
      const int L = s_alist[e];
      if (L >= 0) place(s_arec[e], L, s_len[L]++);
    };
    for (int ll = 0; ll < nlist; ll++) {
      const int lo = s_lo[ll], r = s_cnt[ll];
      if (r == 0) continue;
      const int n0 = s_len[ll], b = n0 - r, base = I.L.base[ll];
      int nlow = 0;
      while (nlow < r && s_sslot[lo + nlow] - base < b) nlow++;
      int topl = n0;
      for (int j = 0; j < nlow; j++) {
        meet(s_sorted[lo + j]);
        for (;;) {
          topl--;
          const int q = topl - b, e = s_tent[lo + q];
          if (e != kNoSlot) {
            meet(e);
            continue;
          }
          place(s_top[lo + q], ll, s_sslot[lo + j] - base);
          break;
        }
      }
      if (topl > b) {  // the remaining top slots all leave: the lowest first, then from the top down
        meet(s_tent[lo]);
        for (int t = topl - 1; t > b; t--) meet(s_tent[lo + (t - b)]);
      }
      s_len[ll] = b;
    }
    s_np = np;
    for (int ll = 0; ll < nlist; ll++) s_pl[ll] = s_len[ll] - s_nf[ll];
  }
  __syncthreads();
  for (int k = tid; k < s_np; k += NT) {
    const int rec = s_prec[k], list = s_plist[k], pos = s_ppos[k];
    if (pos >= I.L.base[list + 1] - I.L.base[list]) {
      atomicOr(&ctl->err, kErrCapacity);
      atomicOr(&gl->halt, 1);
      continue;
    }
    B.lpos[rec] = pos;
    inv[I.L.base[list] + pos] = rec;
  }
}

// One workgroup per simulation of the batch (blockIdx.x).
template <bool kF32, int kSurf>
__global__ __launch_bounds__(kResolveThreads) void k_resolve(const Inst* __restrict__ insts, Glob* gl, Recs RA,
                                                             Recs RB, int* __restrict__ inv, Binned bn, int nscan,
                                                             uint64_t* __restrict__ bestkey, int* __restrict__ killed,
                                                             int reccap) {
  const Inst& I = insts[blockIdx.x];
  if (gl->halt) return;  // a simulation that starts after another one halted skips its step
  __shared__ uint64_t lk[kLdsSort];
  __shared__ int lv[kLdsSort];
  __shared__ int s_scan[kResolveThreads / 64];
  __shared__ int s_wcnt[(kResolveThreads / 64) * kMaxList];
  __shared__ int s_dcount[kMaxSpecies];
  __shared__ int s_ev[kLdsEvents];
  __shared__ int s_kill[kLdsKills];
  __shared__ int s_flag, s_nacc, s_prodoff;
  __shared__ int s_nl[kMaxList], s_nf[kMaxList], s_k0[kMaxList], s_k1[kMaxList], s_pl[kMaxList];
  __shared__ int s_zpre[kMaxRxn0 + 1];  // zeroth-order events before each reaction
  // products adsorbed by their reborn surface check: (slot, record, bound list)
  __shared__ int s_nads;
  __shared__ int s_aslot[kSurf ? kMaxAds : 1], s_arec[kSurf ? kMaxAds : 1], s_alist[kSurf ? kMaxAds : 1];
  const int tid = threadIdx.x;
  const int NT = blockDim.x;
  Control* ctl = I.ctl;
  const DevTables* tab = I.tab;
  Recs B = RB;
  int E = ctl->nedges;
  if (E > I.maxedges) E = I.maxedges;
  int U = ctl->nuni;
  if (U > I.maxevents) U = I.maxevents;
  uint32_t step = ctl->step;
  int nbinned = bn.bstart[nscan - 1];

  // ---- greedy matching: accept an edge when it is the smallest live key at
  // both of its molecules; this is the lexicographically-first maximal
  // matching, i.e. what the sequential loop of bireact accepts.
  if (tid == 0) {
    s_nacc = 0;
    s_nads = 0;
  }
  if (tid < kMaxSpecies) s_dcount[tid] = 0;
  int rounds = 0;
  if (E <= kLdsMatch) {
    // the same rounds in LDS: the edges' molecules hashed to LDS slots
    // (open addressing on the record index), each slot with the smallest
    // key offered to it this round and whether an accepted edge took it;
    // one thread per edge keeps its edge in registers
    uint64_t* best = lk;  // kMatchHash slots (the sort space, unused until the event sort)
    int* srec = lv;
    int* taken = s_kill;
    for (int h = tid; h < kMatchHash; h += NT) {
      srec[h] = -1;
      taken[h] = 0;
      best[h] = kNoKey;
    }
    __syncthreads();
    Edge ed;
    int h1 = 0, h2 = 0, st = tid < E ? 0 : 2;  // 0 open, 1 accepted, 2 out
    auto slot = [&](int rec) {
      int h = (int)(((uint32_t)rec * 2654435761u) >> 20) & (kMatchHash - 1);
      for (;;) {
        const int prev = atomicCAS(&srec[h], -1, rec);
        if (prev == -1 || prev == rec) return h;
        h = (h + 1) & (kMatchHash - 1);
      }
    };
    if (st == 0) {
      ed = I.edges[tid];
      h1 = slot(ed.s1);
      h2 = slot(ed.s2);
    }
    __syncthreads();
    for (;;) {
      if (tid == 0) s_flag = 0;
      __syncthreads();
      if (st == 0) {
        if (taken[h1] || taken[h2]) {
          st = 2;
        } else {
          atomicMin((unsigned long long*)&best[h1], (unsigned long long)ed.key);
          atomicMin((unsigned long long*)&best[h2], (unsigned long long)ed.key);
          s_flag = 1;
        }
      }
      __syncthreads();
      if (!s_flag) break;
      rounds++;
      if (st == 0 && best[h1] == ed.key && best[h2] == ed.key) {
        st = 1;
        taken[h1] = 1;
        taken[h2] = 1;
        I.accepted[atomicAdd(&s_nacc, 1)] = tid;
      }
      __syncthreads();
      if (st == 0) {
        best[h1] = kNoKey;
        best[h2] = kNoKey;
      }
      __syncthreads();
    }
  } else {
  for (int e = tid; e < E; e += NT) I.estate[e] = 0;
  __syncthreads();
  for (;;) {
    if (tid == 0) s_flag = 0;
    __syncthreads();
    for (int e = tid; e < E; e += NT) {
      if (I.estate[e] != 0) continue;
      Edge ed = I.edges[e];
      if (atomic_load_i32(&killed[ed.s1]) || atomic_load_i32(&killed[ed.s2])) {
        I.estate[e] = 2;
        continue;
      }
      atomicMin((unsigned long long*)&bestkey[ed.s1], (unsigned long long)ed.key);
      atomicMin((unsigned long long*)&bestkey[ed.s2], (unsigned long long)ed.key);
      s_flag = 1;
    }
    __syncthreads();
    if (!s_flag) break;
    rounds++;
    for (int e = tid; e < E; e += NT) {
      if (I.estate[e] != 0) continue;
      Edge ed = I.edges[e];
      if (atomic_load_u64(&bestkey[ed.s1]) == ed.key && atomic_load_u64(&bestkey[ed.s2]) == ed.key) {
        I.estate[e] = 1;
        atomicExch(&killed[ed.s1], 1);
        atomicExch(&killed[ed.s2], 1);
        int k = atomicAdd(&s_nacc, 1);
        I.accepted[k] = e;
      }
    }
    __syncthreads();
    for (int e = tid; e < E; e += NT) {
      if (I.estate[e] == 2) continue;
      Edge ed = I.edges[e];
      atomicExch((unsigned long long*)&bestkey[ed.s1], (unsigned long long)kNoKey);
      atomicExch((unsigned long long*)&bestkey[ed.s2], (unsigned long long)kNoKey);
    }
    __syncthreads();
  }
  }
  int B2 = s_nacc;
  __syncthreads();

  // ---- zeroth-order events (zeroreact, smolreact.c:2905-2993): the Poisson
  // count of each reaction is a sum of chunk counts, chunks spread over the
  // threads; the events come first in creation order
  const int nz = tab->nrxn0;
  int Z = 0;
  if (nz > 0) {  // uniform over the workgroup
    if (tid <= kMaxRxn0) s_zpre[tid] = 0;
    __syncthreads();
    for (int r = 0; r < nz; r++) {
      int c = 0;
      for (int j = tid; j < tab->zm[r]; j += NT)
        c += smolrng_poisson_chunk(I.key0, I.key1, (uint32_t)r, (uint32_t)j, step, tab->zg[r]);
      if (c) atomicAdd(&s_zpre[r + 1], c);
    }
    __syncthreads();
    if (tid == 0)
      for (int r = 0; r < nz; r++) s_zpre[r + 1] += s_zpre[r];
    __syncthreads();
    Z = s_zpre[nz];
  }

  // ---- order the events: first-order reactions in live-list order, then
  // bimolecular ones in edge-key order
  if (U + pow2ceil(B2) > I.sortcap || pow2ceil(U + 2 * B2) > I.sortcap || Z + U + B2 > I.sortcap) {
    // more events than the sort space: flag the error, halt the later steps,
    // and finish this step without its reactions so the bookkeeping below
    // (record count, step counter) stays consistent
    if (tid == 0) {
      atomicOr(&ctl->err, kErrSortOverflow);
      atomicOr(&gl->halt, 1);
    }
    for (int i = tid; i < B2; i += NT) {
      Edge ed = I.edges[I.accepted[i]];
      killed[ed.s1] = 0;
      killed[ed.s2] = 0;
    }
    U = 0;
    B2 = 0;
    Z = 0;
    __syncthreads();
  }
  const bool ev_lds = (U + B2) <= kLdsEvents;
  if (U + B2 <= kRankSort) {  // both groups sorted at once, keys gathered straight into LDS
    for (int i = tid; i < U; i += NT) {
      lk[i] = (uint64_t)uni_slot(I, RA, I.uniev[i].rec);
      lv[i] = i;
    }
    for (int i = tid; i < B2; i += NT) {
      const int a = I.accepted[i];
      lk[U + i] = I.edges[a].key;
      lv[U + i] = a;
    }
    __syncthreads();
    wg_rank_sort2<uint64_t>(lk, lv, U, B2, s_ev, nullptr);
  } else {
    for (int i = tid; i < U; i += NT) {
      I.sk[i] = (uint64_t)uni_slot(I, RA, I.uniev[i].rec);
      I.sv[i] = i;
    }
    __syncthreads();
    wg_sort_any(I.sk, I.sv, U, lk, lv, nullptr, ev_lds ? s_ev : nullptr);
    for (int i = tid; i < B2; i += NT) {
      I.sk[U + i] = I.edges[I.accepted[i]].key;
      I.sv[U + i] = I.accepted[i];
    }
    __syncthreads();
    wg_sort_any(I.sk + U, I.sv + U, B2, lk, lv, nullptr, ev_lds ? s_ev + U : nullptr);
  }
  int NE = Z + U + B2;
  const int* evs = ev_lds ? s_ev : I.sv;  // event order: sorted event indices
  // reaction of zeroth-order event e (e < Z)
  auto zrxn = [&](int e) {
    int r = 0;
    while (r + 1 < nz && s_zpre[r + 1] <= e) r++;
    return r;
  };

  // ---- product counts and serial numbers in creation order (each thread
  // takes a contiguous chunk of events so the scan follows creation order)
  int per = (NE + NT - 1) / NT;
  int e0 = tid * per, e1 = min(NE, e0 + per);
  int myprod = 0;
  for (int e = e0; e < e1; e++) {
    const int ee = e - Z;
    const DevRxn& rx = ee < 0    ? tab->rxn0[zrxn(e)]
                       : ee < U ? tab->rxn1[I.uniev[evs[ee]].r]
                                : tab->rxn2[I.edges[evs[ee]].r];
    myprod += rx.nprod;
  }
  int K;
  int pbase = wg_scan(myprod, s_scan, &K);
  // product records of this simulation: [nbinned + s_prodoff, + K)
  if (tid == 0) s_prodoff = atomicAdd(&gl->nprod[step & 1], K);
  int D = U + 2 * B2;  // reactants killed = molsort's kill count
  // few kills: their list slots go straight to LDS for the rank sort
  const bool kill_small = D <= kRankSort;
  int* const kslot = kill_small ? lv : nullptr;
  uint64_t ser0 = ctl->serno_next;

  // ---- products (positions, species, lists) and reactant kills
  int pc = pbase;
  for (int e = e0; e < e1; e++) {
    double rxnpos[3];
    double pAv[3] = {0, 0, 0}, pBv[3] = {0, 0, 0};  // reactant places, in the reaction's reactant order
    double bq[2][3] = {{0, 0, 0}, {0, 0, 0}};       // bounce: places of products 0 and 1
    const DevRxn* rx;
    uint64_t sA = 0, sB = 0;
    int order, ridx;
    const int ee = e - Z;
    if (ee < 0) {  // zeroth order: the product at a uniform position in the open system box
      ridx = zrxn(e);
      rx = &tab->rxn0[ridx];
      order = 0;
      const smolrng4 b = smolrng_philox(I.key0, I.key1, (uint32_t)ridx, (uint32_t)(e - s_zpre[ridx]), step,
                                        SMOLRNG_ZEROPOS);
#pragma unroll
      for (int d = 0; d < 3; d++) rxnpos[d] = tab->zlo[d] + (tab->zhi[d] - tab->zlo[d]) * smolrng_u01open(b.v[d]);
    } else if (ee < U) {
      UniEvent ev = I.uniev[evs[ee]];
      rx = &tab->rxn1[ev.r];
      ridx = ev.r;
      order = 1;
      load_pos<kF32>(RA, I.g, ev.rec, ev.box, rxnpos);
      sA = RA.serno[ev.rec];
      const int us = uni_slot(I, RA, ev.rec);
      if (kill_small) kslot[ee] = us;
      else I.killslot[ee] = us;
    } else {
      Edge ed = I.edges[evs[ee]];
      rx = &tab->rxn2[ed.r];
      ridx = ed.r;
      order = 2;
      double p1[3], p2[3];
      load_pos<kF32>(B, I.g, ed.s1, kF32 ? B.box[ed.s1] : 0, p1);
      load_pos<kF32>(B, I.g, ed.s2, kF32 ? B.box[ed.s2] : 0, p2);
      int id1 = B.ident[ed.s1], id2 = B.ident[ed.s2];
      // bounce by reflection: the reactants' positions before this step's
      // diffusion (posx), by their source records
      const bool refl = rx->bounce && rx->bparam == -2.0 && I.qpos;
      double x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0};
      if (refl) {
        const long q1 = I.src[ed.s1], q2 = I.src[ed.s2];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          x1[d] = I.qpos[3 * q1 + d];
          x2[d] = I.qpos[3 * q2 + d];
        }
      }
      if (ed.wrap && !rx->confspread) {  // morebireact: move the faster molecule next to its partner
        double v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          int cc = (ed.wrap >> (2 * d)) & 3;
          if (cc == 0) v[d] = p2[d] - p1[d];
          else if (cc == 1) v[d] = p2[d] - p1[d] - I.g.width[d];
          else v[d] = p2[d] - p1[d] + I.g.width[d];
        }
        if (tab->difc[id1] < tab->difc[id2]) {
#pragma unroll
          for (int d = 0; d < 3; d++) {
            double shift = p1[d] + v[d] - p2[d];
            p2[d] += shift;
            x2[d] += shift;  // posx moves with the molecule
          }
        } else {
#pragma unroll
          for (int d = 0; d < 3; d++) {
            double shift = p2[d] - v[d] - p1[d];
            p1[d] += shift;
            x1[d] += shift;
          }
        }
      }
      bool swap = (id1 != rx->rct0);
      double* pA = swap ? p2 : p1;
      double* pB = swap ? p1 : p2;
      int iA = swap ? id2 : id1, iB = swap ? id1 : id2;
      double dc1 = tab->difc[iA], dc2 = tab->difc[iB];
      double x = (dc1 == 0 && dc2 == 0) ? 0.5 : dc2 / (dc1 + dc2);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        rxnpos[d] = x * pA[d] + (1.0 - x) * pB[d];
        pAv[d] = pA[d];
        pBv[d] = pB[d];
      }
      sA = B.serno[swap ? ed.s2 : ed.s1];
      sB = B.serno[swap ? ed.s1 : ed.s2];
      bool reflected = false;
      if (refl) {  // doreact's reflection method (smolreact.c:2756-2762)
        double v1[3], v2[3];
        const double pv = sphere_reflect_sphere(swap ? x2 : x1, pAv, swap ? x1 : x2, pBv, rx->bindrad2, v1, v2);
        if (pv > 0 && pv < 1) {
          reflected = true;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            bq[0][d] = v1[d];
            bq[1][d] = v2[d];
          }
        }
      }
      if (rx->bounce && !reflected) {
        // doreact's bounce scaling methods (smolreact.c:2766-2779): move apart along
        // the line joining the reactants
        double bv[3], dist = 0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          bv[d] = pBv[d] - pAv[d];
          dist += bv[d] * bv[d];
        }
        if (dist == 0) {
          dist = sqrt(rx->bindrad2);
          bv[0] = dist;
        } else {
          dist = sqrt(dist);
        }
        const double bx = rx->bparam >= 0 ? 1.0 / dist - (1.0 / rx->bparam)
                                          : 2.0 / dist - 2.0 / (rx->prdpos[0] + rx->prdpos[1]);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          bq[0][d] = pAv[d] - bx * rx->prdpos[0] * bv[d];
          bq[1][d] = pBv[d] + bx * rx->prdpos[1] * bv[d];
        }
      }
      int k2 = U + 2 * (ee - U);
      const int ks1 = I.L.base[tab->listof[id1]] + B.lpos[ed.s1];
      const int ks2 = I.L.base[tab->listof[id2]] + B.lpos[ed.s2];
      if (kill_small) {
        kslot[k2] = ks1;
        kslot[k2 + 1] = ks2;
      } else {
        I.killslot[k2] = ks1;
        I.killslot[k2 + 1] = ks2;
      }
    }
    if (order >= 1) atomicSub(&s_dcount[rx->rct0], 1);
    if (order == 2) atomicSub(&s_dcount[rx->rct1], 1);
    double uvec[3] = {0, 0, 0};
    double rot[9];
    if (rx->offrot) smolrng_rotation(I.key0, I.key1, sA, step, rot);
    else if (rx->rotate) smolrng_unitvec(I.key0, I.key1, sA, step, 3, uvec);
    for (int q = 0; q < rx->nprod; q++) {
      Prod pr;
      pr.ident = rx->prd[q];
      pr.list = rx->prdlist[q];
      pr.sb = 0;
      pr.pad = 0;
      pr.serno = 0;
      double pp[3] = {rxnpos[0], rxnpos[1], rxnpos[2]};
      const double* px0 = rxnpos;  // where the product was made (posx), start of its surface segment
      if (rx->bounce && order == 2) {
        if (q < 2) {  // the reactant moved: same serial, segment from its place
          pp[0] = bq[q][0];
          pp[1] = bq[q][1];
          pp[2] = bq[q][2];
          pr.serno = q == 0 ? sA : sB;
          px0 = q == 0 ? pAv : pBv;
        }
      } else if (rx->confspread && order == 2) {
        // doreact (smolreact.c:2736-2751): products 0 and 1 replace the reactants
        if (q < 2) {
          const double* src = q == 0 ? pAv : pBv;
          pp[0] = src[0];
          pp[1] = src[1];
          pp[2] = src[2];
          pr.serno = q == 0 ? sA : sB;
          px0 = src;
        }
      } else if (rx->fixed) {  // fixed placement: offsets are not rotated (smolreact.c:2811-2812)
        pp[0] = rxnpos[0] + rx->prdvec[q][0];
        pp[1] = rxnpos[1] + rx->prdvec[q][1];
        pp[2] = rxnpos[2] + rx->prdvec[q][2];
      } else if (rx->offrot) {  // offsets off the x axis: one random rotation of the event
        double v1[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          v1[d] = 0;
#pragma unroll
          for (int j = 0; j < 3; j++) v1[d] += rot[3 * d + j] * rx->prdvec[q][j];
        }
        pp[0] = rxnpos[0] + v1[0];
        pp[1] = rxnpos[1] + v1[1];
        pp[2] = rxnpos[2] + v1[2];
      } else if (rx->prdpos[q] != 0) {
        double v1[3] = {rx->prdpos[q] * uvec[0], rx->prdpos[q] * uvec[1], rx->prdpos[q] * uvec[2]};
        pp[0] = rxnpos[0] + v1[0];
        pp[1] = rxnpos[1] + v1[1];
        pp[2] = rxnpos[2] + v1[2];
      }
      if (!kSurf) {
        apply_walls(I.g, pp);  // checkwalls for reborn molecules
        atomicAdd(&s_dcount[pr.ident], 1);
      } else {  // checksurfaces for reborn molecules runs below, one thread per product
        pr.sx = px0[0];
        pr.sy = px0[1];
        pr.sz = px0[2];
      }
      pr.x = pp[0];
      pr.y = pp[1];
      pr.z = pp[2];
      I.prods[pc] = pr;
      pc++;
    }
    LogRow row;
    row.order = order;
    row.r = ridx;
    row.rct[0] = sA;
    row.rct[1] = sB;
    int pfirst = pc - rx->nprod;
    const bool cs = (rx->confspread || rx->bounce) && order == 2;
#pragma unroll
    for (int q = 0; q < kMaxProd; q++)
      row.prd[q] = q < rx->nprod ? (cs && q < 2 ? (q == 0 ? sA : sB) : ser0 + (uint64_t)(pfirst + q)) : 0;
    row.pos[0] = rxnpos[0];
    row.pos[1] = rxnpos[1];
    row.pos[2] = rxnpos[2];
    I.log[e] = row;
  }
  __syncthreads();
  if (kSurf) {
    // checksurfaces for reborn molecules: segment from where each product was
    // made to its place (draws keyed by the product's serial number)
    for (int c = tid; c < K; c += NT) {
      Prod pr = I.prods[c];
      double pp[3] = {pr.x, pr.y, pr.z};
      const double px0[3] = {pr.sx, pr.sy, pr.sz};
      int i2 = pr.ident, bnd = 0;
      // draws keyed by the product's serial number (a reactant's, for
      // products that keep them)
      const uint64_t pser = pr.serno ? pr.serno : ser0 + (uint64_t)c;
      // without panels: the wall pass after the reactions (absorbing walls)
      const int st = I.sf.walls ? (walls_absorb(I.g, pp, px0, tab->difstep[i2], pser, step, I.key0, I.key1, 1)
                                       ? kSrfAbsorbed
                                       : kSrfOk)
                                : srf_check<(kSurf > 1)>(I.sf, I.g, tab->difc, pp, px0, &i2, pser, step,
                                                         SMOLRNG_SURF_REBORN, I.key0, I.key1, &bnd);
      if (st == kSrfAdsorbed && I.sf.blist[i2][bstate(bnd)] >= 0) {
        // bound; it leaves its solution list in the list sort that ends the
        // step (below); bit 2 marks it until then
        pr.sb = bnd | 4;
        atomicAdd((bnd & 3) == 1 ? &ctl->countf[i2] : &ctl->countk[i2], 1);
      } else if (st == kSrfAbsorbed) {
        // molkill: the list sort that ends the step removes it (below)
        pr.sb = kProdKilled;
      } else if (st != kSrfOk || tab->listof[i2] != pr.list) {
        atomicOr(&ctl->err, kErrSurface);
        atomicOr(&gl->halt, 1);
      }
      pr.ident = i2;
      pr.x = pp[0];
      pr.y = pp[1];
      pr.z = pp[2];
      if (pr.sb != kProdKilled) atomicAdd(&s_dcount[pr.ident], 1);
      I.prods[c] = pr;
    }
    __syncthreads();
  }

  // ---- molsort: sort the killed list slots, fill holes from the list tails
  // (swap-from-tail compaction), then append products in the dead-list order
  // the reference produces: creation order reversed, rotated by the kills.
  const bool kill_lds = D <= kLdsKills;
  if (kill_small) {
    // the sort space lk is free here
    wg_bucket_sort(lv, D, s_kill, reinterpret_cast<int*>(lk));
  } else {
    for (int i = tid; i < D; i += NT) {
      I.sk[i] = (uint64_t)I.killslot[i];
      I.sv[i] = 0;
    }
    __syncthreads();
    wg_sort_any(I.sk, I.sv, D, lk, lv, kill_lds ? s_kill : nullptr, nullptr);
  }
  const KillView kv{kill_lds ? s_kill : nullptr, I.sk};
  if (tid < I.L.nlist) {
    int ll = tid;
    int lo = I.L.base[ll], hi = I.L.base[ll + 1];
    int l = 0, h = D;
    while (l < h) {
      int mid = (l + h) >> 1;
      if (kv[mid] < lo) l = mid + 1;
      else h = mid;
    }
    int k0 = l;
    h = D;
    while (l < h) {
      int mid = (l + h) >> 1;
      if (kv[mid] < hi) l = mid + 1;
      else h = mid;
    }
    s_k0[ll] = k0;
    s_k1[ll] = l;
    s_nl[ll] = ctl->nl[ll];
    s_nf[ll] = ctl->nl[ll] - (l - k0);
  }
  __syncthreads();
  // bimolecular reactants stay in the destination buffer as dead records
  for (int i = tid; i < B2; i += NT) {
    Edge ed = I.edges[I.accepted[i]];
    B.ident[ed.s1] = 0;
    B.ident[ed.s2] = 0;
    killed[ed.s1] = 0;
    killed[ed.s2] = 0;
  }
  for (int ll = 0; ll < I.L.nlist; ll++) {
    int k0 = s_k0[ll], k1 = s_k1[ll], nk = k1 - k0;
    int nlc = s_nl[ll], nf = s_nf[ll];
    int lo = I.L.base[ll];
    for (int q = tid; q < nk; q += NT) {
      int t = lo + nf + q;  // tail slot
      int l = k0, h = k1;
      while (l < h) {
        int mid = (l + h) >> 1;
        if (kv[mid] < t) l = mid + 1;
        else h = mid;
      }
      if (l < k1 && kv[l] == t) continue;  // killed
      int above = k1 - l;                          // kills above t
      int rank = (nlc - 1 - (nf + q)) - above;
      int dst = kv[k0 + rank];
      int rec = inv[t];
      B.lpos[rec] = dst - lo;
      inv[dst] = rec;
    }
  }
  __syncthreads();
  // product append order: Q[j] = creation index K-1-((j+D) mod K)
  if (tid < I.L.nlist) s_pl[tid] = 0;
  __syncthreads();
  int overflow = 0;
  const int lane = tid & 63, wv = tid >> 6, nwv = NT >> 6;
  const uint64_t lanes_below = (1ull << lane) - 1;
  for (int j0 = 0; j0 < K; j0 += NT) {
    int j = j0 + tid;
    int c = -1, ll = -1;
    if (j < K) {
      c = K - 1 - (int)(((long)j + D) % K);
      ll = I.prods[c].list;
    }
    // rank of this product among the chunk's products of its list, in j order
    int inwave = 0;
    for (int l2 = 0; l2 < I.L.nlist; l2++) {
      uint64_t m = __ballot(ll == l2);
      if (lane == 0) s_wcnt[wv * kMaxList + l2] = __popcll(m);
      if (ll == l2) inwave = __popcll(m & lanes_below);
    }
    __syncthreads();
    if (ll >= 0) {
      int pre = inwave;
      for (int w2 = 0; w2 < wv; w2++) pre += s_wcnt[w2 * kMaxList + ll];
      int slot = I.L.base[ll] + s_nf[ll] + s_pl[ll] + pre;
      int rec = nbinned + s_prodoff + c;
      if (slot < I.L.base[ll + 1] && rec < reccap) {
        Prod pr = I.prods[c];
        const int pbx = pos2box(I.g, pr.x, pr.y, pr.z);
        const double pp[3] = {pr.x, pr.y, pr.z};
        store_pos<kF32>(B, I.g, rec, pbx, pp);
        const bool pkilled = kSurf && pr.sb == kProdKilled;
        B.ident[rec] = pkilled ? 0 : pr.ident;  // an absorbed product stays as a dead record
        if (B.sb) B.sb[rec] = pkilled ? 0 : pr.sb & ~4;
        if (kSurf && (pkilled || (pr.sb & 4))) {
          const int a = atomicAdd(&s_nads, 1);
          if (a < kMaxAds) {
            s_aslot[a] = slot;
            s_arec[a] = rec;
            s_alist[a] = pkilled ? -1 : I.sf.blist[pr.ident][bstate(pr.sb)];
          }
        }
        if (B.inst) B.inst[rec] = blockIdx.x;  // batch member (multi-simulation records)
        B.serno[rec] = pr.serno ? pr.serno : ser0 + (uint64_t)c;
        B.lpos[rec] = slot - I.L.base[ll];
        B.box[rec] = pbx;
        inv[slot] = rec;
      } else {
        overflow = 1;
        if (rec < reccap) B.ident[rec] = 0;  // no molecule: the next step skips the record
      }
    }
    __syncthreads();
    if (tid < I.L.nlist) {
      int t = 0;
      for (int w2 = 0; w2 < nwv; w2++) t += s_wcnt[w2 * kMaxList + tid];
      s_pl[tid] += t;
    }
    __syncthreads();
  }
  if (overflow) atomicOr(&ctl->err, kErrCapacity);
  if (kSurf && I.sf.walls) {
    // survivors the second wall pass kills (k_diffuse drew their coins): the
    // ones that did not react die here and leave in the sort below
    __syncthreads();
    const int nq = min(ctl->nwall, I.maxevents);
    for (int q = tid; q < nq; q += NT) {
      const int rec = I.sf.wq[q];
      if (rec < 0) continue;
      const int i = B.ident[rec];
      if (i <= 0) continue;  // a bimolecular reactant of this step
      const int a = atomicAdd(&s_nads, 1);
      if (a < kMaxAds) {
        s_aslot[a] = I.L.base[tab->listof[i]] + B.lpos[rec];
        s_arec[a] = rec;
        s_alist[a] = -1;
      }
      B.ident[rec] = 0;
      atomicSub(&s_dcount[i], 1);
    }
  }
  if (kSurf) {
    __syncthreads();
    end_of_step_sort(I, ctl, gl, B, inv, s_nads, s_aslot, s_arec, s_alist, s_nf, s_pl, lv, s_kill, (int*)lk);
  }
  // ---- species counts (accumulated in LDS by the product loop), list
  // lengths, counters
  __syncthreads();
  if (tid < kMaxSpecies && s_dcount[tid]) ctl->count[tid] += s_dcount[tid];
  if (tid == 0) {
    int run = 0;
    unsigned long long live0 = 0;
    for (int ll = 0; ll < I.L.nlist; ll++) live0 += (unsigned long long)s_nl[ll];
    ctl->molsteps += live0;
    for (int ll = 0; ll < I.L.nlist; ll++) {
      int nlnew = s_nf[ll] + s_pl[ll];
      int cap = I.L.base[ll + 1] - I.L.base[ll];
      if (nlnew > cap) nlnew = cap;
      ctl->nl[ll] = nlnew;
      ctl->nfresh[ll] = 0;
      ctl->pfx[ll] = run;
      if (tab->diffuselist[ll]) run += nlnew;
    }
    int top = nbinned + s_prodoff + K;
    atomicMax(&gl->nrec[step & 1], top < reccap ? top : reccap);
    if (blockIdx.x == 0) {  // counters of the next step
      gl->nrec[(step + 1) & 1] = 0;
      gl->nprod[(step + 1) & 1] = 0;
      gl->ncand = 0;
      gl->par = step & 1;
    }
    ctl->nbinned = nbinned;
    ctl->serno_next = ser0 + (uint64_t)K;
    ctl->naccept = B2;
    ctl->nlog = NE;
    if (E > ctl->maxedges_seen) ctl->maxedges_seen = E;
    ctl->rounds = rounds;
    ctl->nedges = 0;
    ctl->nuni = 0;
    ctl->nwall = 0;
    ctl->step = step + 1;
  }
}

}  // namespace smolamd

// ===========================================================================
// host side

namespace smolamd {

template <typename T>
static int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return -1;
  return 0;
}

static int alloc_recs(Recs* r, size_t n, bool multi, bool f32, bool surf) {
  int ok = 0;
  if (surf) ok |= dalloc(&r->sb, n) | dalloc(&r->ov, n);
  if (f32) ok |= dalloc(&r->ox, n) | dalloc(&r->oy, n) | dalloc(&r->oz, n);
  else ok |= dalloc(&r->px, n) | dalloc(&r->py, n) | dalloc(&r->pz, n);
  ok |= dalloc(&r->ident, n) | dalloc(&r->serno, n) | dalloc(&r->lpos, n) | dalloc(&r->box, n) |
        dalloc(&r->rank, n);
  if (multi) ok |= dalloc(&r->inst, n);
  return ok;
}

static void free_recs(Recs* r) {
  void* ptrs[] = {r->px, r->py, r->pz, r->ox, r->oy, r->oz, r->inst, r->ident, r->serno, r->lpos, r->box, r->rank,
                  r->sb, r->ov};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

static int upload_instances(DeviceState* s) {
  HIPOK(hipMemcpy(s->dinst, s->hinst.data(), sizeof(Inst) * s->ninst, hipMemcpyHostToDevice));
  return 0;
}

// Species rows of one simulation's tables (the fields k_diffuse reads per
// molecule; species without a list or reactions get harmless rows).
static void spec_rows(const DevTables& t, SpecRow* out) {
  for (int i = 0; i < kMaxSpecies; i++) {
    SpecRow& r = out[i];
    memset(&r, 0, sizeof(r));
    r.ds = t.difstep[i];
    r.ll = t.listof[i];
    r.nu = t.nuni[i];
    for (int j = 0; j < kMaxRxnPerSet; j++) {
      const int x = t.uni[i][j];
      const bool ok = j < r.nu && x >= 0 && x < kMaxRxn;
      r.rx[j] = ok ? x : -1;
      r.prob[j] = ok && t.rxn1[x].permit_soln ? t.rxn1[x].prob : -1.0;
    }
  }
}

DeviceState* engine_create_batch(int n, const EngineConfig* cfg, const DevTables* tab, const Geometry* geo,
                                 const uint32_t* key0, const uint32_t* key1, const double* gtable, int ngauss,
                                 const HostSurf* surf, char* err, int errlen) {
  DeviceState* s = new DeviceState();
  s->ninst = n;
  s->multi = n > 1;
  s->f32 = cfg[0].f32 != 0;
  s->ngauss = ngauss;
  for (int k = 1; k < n; k++)
    if ((cfg[k].f32 != 0) != s->f32) {
      snprintf(err, errlen, "ensemble members must use the same record precision");
      delete s;
      return nullptr;
    }
  s->cfg.assign(cfg, cfg + n);
  s->htab.assign(tab, tab + n);
  s->hinst.assign(n, Inst());
  if (s->multi && surf && surf->on) {
    snprintf(err, errlen, "surfaces are only supported for a single simulation");
    delete s;
    return nullptr;
  }
  // slot, key and per-simulation array layout
  int run = 0, keys = 0;
  size_t nev = 0, ned = 0, nsort = 0;
  std::vector<size_t> evoff(n), edoff(n), sortoff(n);
  long prodroom = 0;
  for (int k = 0; k < n; k++) {
    Inst& I = s->hinst[k];
    memset(&I, 0, sizeof(Inst));
    I.L.nlist = cfg[k].nlist;
    for (int ll = 0; ll < cfg[k].nlist; ll++) {
      I.L.base[ll] = run;
      run += cfg[k].cap[ll];
    }
    for (int ll = cfg[k].nlist; ll <= kMaxList; ll++) I.L.base[ll] = run;
    I.g = geo[k];
    I.key0 = key0[k];
    I.key1 = key1[k];
    // z sub-bins: the largest power of two whose slab height is at least the
    // largest screening radius (binding radius + FP32 margin)
    {
      double rmax = 0;
      for (int i = 0; i < kMaxSpecies; i++)
        if (tab[k].reactmask[i]) rmax = fmax(rmax, sqrt(tab[k].maxbind2[i]) + geo[k].prefilter_margin);
      // at most 4, and not below about one molecule per sub-bin
      int zs = 1;
      while (zs < 4 && 2 * zs <= cfg[k].expected_per_box + 0.5 && rmax > 0 &&
             geo[k].bsize[2] / (2 * zs) >= rmax * (1 + 1e-9))
        zs *= 2;
      if (rmax <= 0) zs = 1;
      I.zsub = cfg[k].zsub_max > 0 && zs > cfg[k].zsub_max ? cfg[k].zsub_max : zs;
    }
    I.keybase = keys;
    if ((long)keys + (long)cfg[k].nlist * geo[k].nbox * I.zsub >= (1L << 30)) {
      snprintf(err, errlen, "too many virtual boxes for the scan");
      delete s;
      return nullptr;
    }
    keys += cfg[k].nlist * geo[k].nbox * I.zsub;
    I.maxevents = cfg[k].maxevents;
    I.maxedges = cfg[k].maxedges;
    I.sortcap = pow2ceil(2 * cfg[k].maxevents + 16);
    evoff[k] = nev;
    edoff[k] = ned;
    sortoff[k] = nsort;
    nev += (size_t)I.maxevents;
    ned += (size_t)I.maxedges;
    nsort += (size_t)I.sortcap;
    // product records one step can create: the single-simulation engine
    // keeps the generous bound of the sort space; batch members use their
    // event capacity (overflow sets kErrCapacity)
    prodroom += n == 1 ? 2L * I.sortcap : 2L * cfg[k].maxevents + 16;
    // pair-screen constants: largest binding radius + FP32 margin
    for (int i = 0; i < kMaxSpecies; i++) {
      I.pc.reactmask[i] = tab[k].reactmask[i];
      double rr = sqrt(tab[k].maxbind2[i]) + geo[k].prefilter_margin;
      I.pc.r2f[i] = (float)(rr * rr);
    }
    double ext = 0;
    for (int d = 0; d < 3; d++) {
      ext = fmax(ext, fabs(geo[k].low[d]));
      ext = fmax(ext, fabs(geo[k].high[d]));
      ext = fmax(ext, geo[k].high[d] - geo[k].low[d]);
    }
    I.screen_lim = 2.0 * ext;
  }
  s->captot = run;
  if ((keys + 1 + kScanTile - 1) / kScanTile > kScanBlock * 8) {
    // too many keys for the scan with z sub-bins: one sub-bin per box
    keys = 0;
    for (int k = 0; k < n; k++) {
      Inst& I = s->hinst[k];
      I.zsub = 1;
      I.keybase = keys;
      keys += cfg[k].nlist * geo[k].nbox;
    }
  }
  // the pair search's half-stencil specialization: in every simulation the
  // only reacting list pair is one list with itself
  s->half1 = true;
  for (int k = 0; k < n; k++) {
    int ndiag = 0;
    for (int l1 = 0; l1 < cfg[k].nlist; l1++)
      for (int l2 = 0; l2 < cfg[k].nlist; l2++)
        if (tab[k].rxnmol2[l1 * kMaxList + l2]) {
          if (l1 != l2) s->half1 = false;
          else ndiag++;
        }
    if (ndiag > 1) s->half1 = false;
  }
  s->nscan = keys + 1;
  s->nscanblk = (s->nscan + kScanTile - 1) / kScanTile;
  if (s->nscanblk > kScanBlock * 8) {
    snprintf(err, errlen, "too many virtual boxes for the scan (%d)", s->nscan);
    delete s;
    return nullptr;
  }
  if ((long)run + prodroom >= (1L << 31)) {
    snprintf(err, errlen, "too many molecule records");
    delete s;
    return nullptr;
  }
  s->reccap = run + (int)prodroom;
  size_t rc = (size_t)s->reccap;
  int ok = 0;
  ok |= alloc_recs(&s->buf[0], rc, s->multi, s->f32, surf && surf->on) |
        alloc_recs(&s->buf[1], rc, s->multi, s->f32, surf && surf->on);
  ok |= dalloc(&s->inv, (size_t)s->captot);
  ok |= dalloc(&s->bn.count, s->nscan) | dalloc(&s->bn.boff, s->nscanblk) |
        dalloc(&s->bn.bstart, s->nscan) | dalloc(&s->bn.spk, rc);
  ok |= dalloc(&s->bestkey, rc) | dalloc(&s->killed, rc);
  ok |= dalloc(&s->edges, ned) | dalloc(&s->estate, ned) | dalloc(&s->accepted, ned);
  // screened candidates: the FP32 screen passes few pairs beyond the ones
  // within the binding radius, so the edge capacity bounds them with room
  s->candcap = ned * 2 > (size_t)(1 << 28) ? (1 << 28) : (int)(ned * 2);
  ok |= dalloc(&s->cand, (size_t)s->candcap);
  ok |= dalloc(&s->uniev, nev) | dalloc(&s->log, nsort);
  ok |= dalloc(&s->sortkey, nsort) | dalloc(&s->sortval, nsort) | dalloc(&s->killslot, nsort);
  ok |= dalloc(&s->prods, nsort * kMaxProd);
  ok |= dalloc(&s->ctl, (size_t)n) | dalloc(&s->tab, (size_t)n) | dalloc(&s->gtab, ngauss);
  ok |= dalloc(&s->srow, (size_t)n * kMaxSpecies);
  ok |= dalloc(&s->dinst, (size_t)n) | dalloc(&s->glob, 1);
  if (surf && surf->on) {
    ok |= dalloc(&s->d_pnl, surf->npnl > 0 ? surf->npnl : 1) | dalloc(&s->d_bpstart, (size_t)surf->nbox + 1) |
          dalloc(&s->d_bplist, surf->bpstart[surf->nbox] > 0 ? surf->bpstart[surf->nbox] : 1) |
          dalloc(&s->d_act, (size_t)surf->nsrf * kMaxSpecies * 2) |
          dalloc(&s->d_actb, (size_t)surf->nsrf * kMaxSpecies * 2) | dalloc(&s->moves, nev) | dalloc(&s->wq, nev);
  }
  // tile binning: one simulation without surfaces (SMOLDYN_AMD_TILE=0 keeps
  // the four-launch counting sort for every step).  Tiles are sized for about
  // 640 records, so most fit the records k_tile2 keeps in registers.
  {
    const char* env = getenv("SMOLDYN_AMD_TILE");
    if (!s->multi && !(surf && surf->on) && !(env && env[0] == '0')) {
      const int nkeys = s->nscan - 1;
      const double per_key = fmax(cfg[0].expected_per_box, 1e-3) / (double)s->hinst[0].zsub;
      int sh = kTileMinShift;
      while (sh < kTileMaxShift && (double)(1 << (sh + 1)) * per_key <= 640.0) sh++;
      auto& T = s->ta;
      T.tshift = sh;
      T.nkeys = nkeys;
      T.ntile = (nkeys + (1 << sh) - 1) >> sh;
      ok |= dalloc(&T.toff, T.ntile) | dalloc(&T.nkey, rc);
      ok |= dalloc(&T.stay, 2 * (size_t)T.ntile) | dalloc(&T.range, 2 * (size_t)T.ntile) |
            dalloc(&T.imm, 2 * (size_t)T.ntile) |
            dalloc(&T.bucket, (size_t)T.ntile * kTileBucket) | dalloc(&T.spill, rc) | dalloc(&T.nspill, 2);
      s->tile_ok = true;
    }
  }
  if (!s->multi && cfg[0].reflect) ok |= dalloc(&s->qpos, 3 * rc) | dalloc(&s->srcmap, rc);
  if (ok) {
    snprintf(err, errlen, "device allocation failed (%d molecule records)", s->reccap);
    engine_destroy(s);
    return nullptr;
  }
  for (int k = 0; k < n; k++) {
    Inst& I = s->hinst[k];
    I.tab = AS_VIEW(const DevTables, s->tab + k);
    I.srow = AS_VIEW(const SpecRow, s->srow + (size_t)k * kMaxSpecies);
    I.ctl = AS_VIEW(Control, s->ctl + k);
    I.uniev = AS_VIEW(UniEvent, s->uniev + evoff[k]);
    I.edges = AS_VIEW(Edge, s->edges + edoff[k]);
    I.estate = AS_VIEW(int, s->estate + edoff[k]);
    I.accepted = AS_VIEW(int, s->accepted + edoff[k]);
    I.log = AS_VIEW(LogRow, s->log + sortoff[k]);
    I.sk = AS_VIEW(uint64_t, s->sortkey + sortoff[k]);
    I.sv = AS_VIEW(int, s->sortval + sortoff[k]);
    I.killslot = AS_VIEW(int, s->killslot + sortoff[k]);
    I.prods = AS_VIEW(Prod, s->prods + sortoff[k] * kMaxProd);
    I.qpos = AS_VIEW(double, s->qpos);
    I.src = AS_VIEW(int, s->srcmap);
  }
  if (surf && surf->on) {
    Inst& I = s->hinst[0];
    if (surf->npnl > 0) (void)hipMemcpy(s->d_pnl, surf->pnl, sizeof(DevPanel) * surf->npnl, hipMemcpyHostToDevice);
    (void)hipMemcpy(s->d_bpstart, surf->bpstart, sizeof(int) * (surf->nbox + 1), hipMemcpyHostToDevice);
    if (surf->bpstart[surf->nbox] > 0)
      (void)hipMemcpy(s->d_bplist, surf->bplist, sizeof(int) * surf->bpstart[surf->nbox], hipMemcpyHostToDevice);
    if (surf->nsrf > 0) {
      (void)hipMemcpy(s->d_act, surf->act, sizeof(DevSrfAct) * surf->nsrf * kMaxSpecies * 2, hipMemcpyHostToDevice);
      (void)hipMemcpy(s->d_actb, surf->actb, sizeof(DevSrfAct) * surf->nsrf * kMaxSpecies * 2, hipMemcpyHostToDevice);
    }
    I.sf.walls = surf->walls;
    I.sf.wq = AS_VIEW(int, s->wq);
    I.sf.on = 1;
    for (int ll = 0; ll < kMaxList; ll++) I.sf.listbound[ll] = surf->listbound[ll];
    for (int i = 0; i < kMaxSpecies; i++) {
      I.sf.blist[i][0] = surf->blist[i][0];
      I.sf.blist[i][1] = surf->blist[i][1];
    }
    I.sf.actb = AS_VIEW(const DevSrfAct, s->d_actb);
    I.moves = AS_VIEW(MoveEv, s->moves);
    I.sf.npnl = surf->npnl;
    for (int ll = 0; ll < kMaxList; ll++) I.sf.listdiffuse[ll] = surf->listdiffuse[ll];
    I.sf.epsilon = surf->epsilon;
    I.sf.pnl = AS_VIEW(const DevPanel, s->d_pnl);
    I.sf.bpstart = AS_VIEW(const int, s->d_bpstart);
    I.sf.bplist = AS_VIEW(const int, s->d_bplist);
    I.sf.act = AS_VIEW(const DevSrfAct, s->d_act);
    s->surf_on = true;
    s->surf_curved = false;
    for (int k = 0; k < surf->npnl; k++)
      if (surf->pnl[k].ps == kPsTri || surf->pnl[k].ps == kPsCyl || surf->pnl[k].ps == kPsHemi) s->surf_curved = true;
  }
  (void)hipMemset(s->bn.count, 0, sizeof(int) * s->nscan);
  if (s->tile_ok) {
    (void)hipMemset(s->ta.imm, 0, sizeof(int) * 2 * s->ta.ntile);
    (void)hipMemset(s->ta.stay, 0, sizeof(int) * 2 * s->ta.ntile);
    (void)hipMemset(s->ta.range, 0, sizeof(int2) * 2 * s->ta.ntile);
    (void)hipMemset(s->ta.nspill, 0, sizeof(int) * 2);
  }
  (void)hipMemset(s->bestkey, 0xFF, sizeof(uint64_t) * rc);
  (void)hipMemset(s->killed, 0, sizeof(int) * rc);
  (void)hipMemset(s->ctl, 0, sizeof(Control) * n);
  {
    Glob g0;
    memset(&g0, 0, sizeof(g0));
    g0.ninst = n;
    g0.par = 1;  // step 0 reads nrec[1]
    (void)hipMemcpy(s->glob, &g0, sizeof(Glob), hipMemcpyHostToDevice);
  }
  (void)hipMemcpy(s->tab, tab, sizeof(DevTables) * n, hipMemcpyHostToDevice);
  {
    std::vector<SpecRow> rows((size_t)n * kMaxSpecies);
    for (int k = 0; k < n; k++) spec_rows(tab[k], &rows[(size_t)k * kMaxSpecies]);
    (void)hipMemcpy(s->srow, rows.data(), sizeof(SpecRow) * rows.size(), hipMemcpyHostToDevice);
  }
  (void)hipMemcpy(s->gtab, gtable, sizeof(double) * ngauss, hipMemcpyHostToDevice);
  if (upload_instances(s)) {
    snprintf(err, errlen, "device copy failed");
    engine_destroy(s);
    return nullptr;
  }
  if (hipStreamCreate(&s->stream) != hipSuccess) {
    snprintf(err, errlen, "stream creation failed");
    engine_destroy(s);
    return nullptr;
  }
  for (int k = 0; k < 16; k++) (void)hipEventCreate(&s->ev[k]);
  for (int k = 0; k < 2 * kTimedSlots; k++) (void)hipEventCreate(&s->kev[k]);
  return s;
}

DeviceState* engine_create(const EngineConfig& cfg, const DevTables& tab, const Geometry& geo, const double* gtable,
                           int ngauss, uint32_t key0, uint32_t key1, const HostSurf* surf, char* err, int errlen) {
  return engine_create_batch(1, &cfg, &tab, &geo, &key0, &key1, gtable, ngauss, surf, err, errlen);
}

void engine_destroy(DeviceState* s) {
  if (!s) return;
  // the stream's work ends, its graphs and events go, then the memory they used
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int k = 0; k < 2; k++)
    if (s->graph_ready[k]) (void)hipGraphExecDestroy(s->graph[k]);
  if (s->stream) {
    for (int k = 0; k < 16; k++)
      if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
    for (int k = 0; k < 2 * kTimedSlots; k++)
      if (s->kev[k]) (void)hipEventDestroy(s->kev[k]);
    (void)hipStreamDestroy(s->stream);
  }
  free_recs(&s->buf[0]);
  free_recs(&s->buf[1]);
  void* ptrs[] = {s->inv,     s->bn.count, s->bn.boff, s->bn.bstart, s->bn.spk, s->edges,  s->cand,  s->estate,
                  s->accepted, s->bestkey, s->killed,  s->uniev,   s->log,       s->sortkey, s->sortval, s->killslot,
                  s->ctl,     s->tab,     s->gtab,    s->inj,     s->prods,     s->d_pnl,   s->d_bpstart,
                  s->d_bplist, s->d_act,  s->d_actb,  s->moves,   s->dinst,   s->glob,    s->obs,
                  s->wq,      s->srow,    s->ta.stay, s->ta.range, s->ta.imm, s->ta.bucket, s->ta.spill,
                  s->ta.nspill, s->ta.toff,   s->ta.nkey, s->qpos,   s->srcmap};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

// Upload list ll of simulation inst in live order.  Records are appended
// after the records uploaded before; the first step bins everything.
int engine_upload_list_inst(DeviceState* s, int inst, int ll, int n, const int* ident, const double* pos,
                            const uint64_t* serno) {
  if (inst < 0 || inst >= s->ninst) return -2;
  const Inst& I = s->hinst[inst];
  if (ll < 0 || ll >= I.L.nlist) return -2;
  int cap = I.L.base[ll + 1] - I.L.base[ll];
  if (n > cap) return -3;
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl + inst, sizeof(Control), hipMemcpyDeviceToHost));
  int r0 = s->hnrec;
  if (r0 + n > s->reccap) return -3;
  Recs& R = s->buf[s->cur];
  if (n) {
    std::vector<double> x(n), y(n), z(n);
    std::vector<int> lp(n), inv(n), box(n, 0), in(n, inst);
    for (int k = 0; k < n; k++) {
      x[k] = pos[3 * k];
      y[k] = pos[3 * k + 1];
      z[k] = pos[3 * k + 2];
      lp[k] = k;
      inv[k] = r0 + k;
    }
    // each molecule's virtual box (pos2box, smolboxes.c:67-79), as a read
    // of the state before the first step reports it
    const Geometry& g = I.g;
    for (int k = 0; k < n; k++) {
      int b = 0;
      for (int d = 0; d < 3; d++) {
        int indx = (int)((pos[3 * k + d] - g.bmin[d]) / g.bsize[d]);
        if (indx < 0) indx = 0;
        else if (indx >= g.side[d]) indx = g.side[d] - 1;
        b = g.side[d] * b + indx;
      }
      box[k] = b;
    }
    if (s->f32) {
      // FP32 offsets from the origin of each molecule's virtual box
      std::vector<float> fx(n), fy(n), fz(n);
      for (int k = 0; k < n; k++) {
        const int b = box[k];
        double o[3];
        box_origin(g, b, o);
        fx[k] = (float)(x[k] - o[0]);
        fy[k] = (float)(y[k] - o[1]);
        fz[k] = (float)(z[k] - o[2]);
      }
      HIPOK(hipMemcpy(R.ox + r0, fx.data(), sizeof(float) * n, hipMemcpyHostToDevice));
      HIPOK(hipMemcpy(R.oy + r0, fy.data(), sizeof(float) * n, hipMemcpyHostToDevice));
      HIPOK(hipMemcpy(R.oz + r0, fz.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    } else {
      HIPOK(hipMemcpy(R.px + r0, x.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      HIPOK(hipMemcpy(R.py + r0, y.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      HIPOK(hipMemcpy(R.pz + r0, z.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    HIPOK(hipMemcpy(R.ident + r0, ident, sizeof(int) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(R.serno + r0, serno, sizeof(uint64_t) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(R.lpos + r0, lp.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(R.box + r0, box.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if (R.sb) {  // uploaded molecules are in solution
      std::vector<int> zero(n, 0);
      HIPOK(hipMemcpy(R.sb + r0, zero.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    if (s->multi) HIPOK(hipMemcpy(R.inst + r0, in.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(s->inv + I.L.base[ll], inv.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  }
  s->hnrec = r0 + n;
  s->sorted = false;
  c.nl[ll] = n;
  int run = 0;
  for (int l2 = 0; l2 < I.L.nlist; l2++) {
    c.pfx[l2] = run;
    if (s->htab[inst].diffuselist[l2]) run += c.nl[l2];
  }
  HIPOK(hipMemcpy(s->ctl + inst, &c, sizeof(Control), hipMemcpyHostToDevice));
  return 0;
}

int engine_upload_list(DeviceState* s, int ll, int n, const int* ident, const double* pos, const uint64_t* serno) {
  return engine_upload_list_inst(s, 0, ll, n, ident, pos, serno);
}

int engine_append_fresh(DeviceState* s, int ll, int n, const int* ident, const double* pos, const uint64_t* serno) {
  if (s->multi || s->surf_on || s->f32) return -1;
  const Inst& I = s->hinst[0];
  if (ll < 0 || ll >= I.L.nlist || n < 0) return -2;
  if (n == 0) return 0;
  HIPOK(hipStreamSynchronize(s->stream));
  Control c;
  Glob g;
  HIPOK(hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(&g, s->glob, sizeof(Glob), hipMemcpyDeviceToHost));
  const int slot = (c.step + 1) & 1;  // the current buffer's record count (cur_nrec)
  const int r0 = g.nrec[slot];
  if (r0 + n > s->reccap || c.nl[ll] + n > I.L.base[ll + 1] - I.L.base[ll]) return -3;
  // the molecules added since the last step: records [fresh_rec0, r0) with
  // their lists; a step since the last addition clears them
  bool any = false;
  for (int l2 = 0; l2 < I.L.nlist; l2++) any |= c.nfresh[l2] > 0;
  if (!any || s->fresh_step != c.step) {
    s->fresh_rec0 = r0;
    s->fresh_list.clear();
    s->fresh_lp.clear();
    s->fresh_step = c.step;
  }
  Recs& R = s->buf[s->cur];
  std::vector<double> x(n), y(n), z(n);
  std::vector<int> box(n, 0);
  for (int k = 0; k < n; k++) {
    x[k] = pos[3 * k];
    y[k] = pos[3 * k + 1];
    z[k] = pos[3 * k + 2];
    s->fresh_list.push_back(ll);
  }
  HIPOK(hipMemcpy(R.px + r0, x.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(R.py + r0, y.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(R.pz + r0, z.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(R.ident + r0, ident, sizeof(int) * n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(R.serno + r0, serno, sizeof(uint64_t) * n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(R.box + r0, box.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  // live positions of list ll's added molecules: the dead list hands them to
  // the molsort latest first, so this call's molecules (already in that
  // order) go before the ones added earlier since the last step
  const int nf = c.nfresh[ll] + n, first = c.nl[ll] + n - nf;
  const int nold = r0 - s->fresh_rec0, nr = nold + n;
  std::vector<int> inv(nf), prev;
  s->fresh_lp.resize(nr);
  for (int k = 0; k < nold; k++)
    if (s->fresh_list[k] == ll) prev.push_back(k);
  std::sort(prev.begin(), prev.end(), [&](int a, int b) { return s->fresh_lp[a] < s->fresh_lp[b]; });
  int q = 0;
  for (int k = nold; k < nr; k++) {  // this call's records
    s->fresh_lp[k] = first + q;
    inv[q++] = s->fresh_rec0 + k;
  }
  for (int k : prev) {  // earlier ones of list ll, shifted behind in their order
    s->fresh_lp[k] = first + q;
    inv[q++] = s->fresh_rec0 + k;
  }
  HIPOK(hipMemcpy(R.lpos + s->fresh_rec0, s->fresh_lp.data(), sizeof(int) * nr, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(s->inv + I.L.base[ll] + first, inv.data(), sizeof(int) * nf, hipMemcpyHostToDevice));
  c.nl[ll] += n;
  c.nfresh[ll] = nf;
  for (int k = 0; k < n; k++) {
    c.count[ident[k]]++;
    if (serno[k] >= c.serno_next) c.serno_next = serno[k] + 1;
  }
  g.nrec[slot] = r0 + n;
  s->sorted = false;  // the next step bins the added records with the counting sort
  HIPOK(hipMemcpy(s->ctl, &c, sizeof(Control), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(s->glob, &g, sizeof(Glob), hipMemcpyHostToDevice));
  return 0;
}

// Counters of simulation inst; also sets the record count of the uploaded
// state (every simulation of a batch starts at the same step).
int engine_set_counters_inst(DeviceState* s, int inst, uint64_t serno_next, uint32_t step, const int* counts) {
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl + inst, sizeof(Control), hipMemcpyDeviceToHost));
  c.serno_next = serno_next;
  c.step = step;
  for (int i = 0; i < kMaxSpecies; i++) c.count[i] = counts ? counts[i] : 0;
  HIPOK(hipMemcpy(s->ctl + inst, &c, sizeof(Control), hipMemcpyHostToDevice));
  Glob g;
  memset(&g, 0, sizeof(g));
  g.nrec[(step + 1) & 1] = s->hnrec;
  g.par = (step + 1) & 1;
  g.ninst = s->ninst;
  HIPOK(hipMemcpy(s->glob, &g, sizeof(Glob), hipMemcpyHostToDevice));
  return 0;
}

int engine_set_counters(DeviceState* s, uint64_t serno_next, uint32_t step, const int* counts) {
  return engine_set_counters_inst(s, 0, serno_next, step, counts);
}

// Launch one step from buffer src into 1-src on s->stream.  Grids cover the
// record capacity; kernels read the exact counts from device memory.
// timed: events between the kernels into ev[0..7].
static void launch_step(DeviceState* s, int src, const double* dinj, int timed, bool tile) {
  int nb = (s->reccap + kBlock - 1) / kBlock;
  hipStream_t st = s->stream;
  Recs A = s->buf[src], B = s->buf[1 - src];
  const bool f = s->f32;
  // surface variants: 2 when the model has triangles, cylinders or
  // hemispheres, 1 for rectangles, spheres and disks only
  const int sv = s->surf_on ? (s->surf_curved ? 2 : 1) : 0;
  auto kdiff = s->multi ? (f ? k_diffuse<0, true, true> : k_diffuse<0, true, false>)
               : sv == 2 ? (f ? k_diffuse<2, false, true> : k_diffuse<2, false, false>)
               : sv == 1 ? (f ? k_diffuse<1, false, true> : k_diffuse<1, false, false>)
                         : (f ? k_diffuse<0, false, true> : k_diffuse<0, false, false>);
  auto kscat = s->multi ? (f ? k_scatter<true, true, false> : k_scatter<true, false, false>)
               : s->surf_on ? (f ? k_scatter<false, true, true> : k_scatter<false, false, true>)
                            : (f ? k_scatter<false, true, false> : k_scatter<false, false, false>);
  auto kpair = s->multi ? (s->half1 ? k_pairs<true, 0, true> : k_pairs<true, 0, false>)
               : s->surf_on ? (s->half1 ? k_pairs<false, 1, true> : k_pairs<false, 1, false>)
                            : (s->half1 ? k_pairs<false, 0, true> : k_pairs<false, 0, false>);
  auto kexact = s->multi ? (f ? k_exact<0, true, true> : k_exact<0, true, false>)
                : sv == 2 ? (f ? k_exact<2, false, true> : k_exact<2, false, false>)
                : sv == 1 ? (f ? k_exact<1, false, true> : k_exact<1, false, false>)
                          : (f ? k_exact<0, false, true> : k_exact<0, false, false>);
  auto kres = sv == 2 ? (f ? k_resolve<true, 2> : k_resolve<false, 2>)
              : sv == 1 ? (f ? k_resolve<true, 1> : k_resolve<false, 1>)
                        : (f ? k_resolve<true, 0> : k_resolve<false, 0>);
  // timed: each kernel carries start / stop events stamped at its own
  // dispatch start and end (the timestamps a kernel trace reads)
  hipEvent_t* kev = s->kev;
  s->timed_mask = 0;
  auto go = [&](int slot, auto kernel, dim3 grid, dim3 block, auto... args) {
    if (timed) {
      hipExtLaunchKernelGGL(kernel, grid, block, 0, st, kev[2 * slot], kev[2 * slot + 1], 0, args...);
      s->timed_mask |= 1 << slot;
    } else {
      hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    }
  };
  if (tile) {
    Tiles tl;
    tl.ntile = s->ta.ntile;
    tl.toff = s->ta.toff;
    tl.nkey = s->ta.nkey;
    tl.reccap = s->reccap;
    tl.nkeys = s->ta.nkeys;
    tl.tshift = s->ta.tshift;
    tl.stay = s->ta.stay;
    tl.range = s->ta.range;
    tl.imm = s->ta.imm;
    tl.bucket = s->ta.bucket;
    tl.spill = s->ta.spill;
    tl.nspill = s->ta.nspill;
    go(kTTile1, f ? k_tile1<true> : k_tile1<false>, dim3((nb + kTile1Per - 1) / kTile1Per), dim3(kBlock), s->dinst,
       s->glob, A,
       (const int*)s->bn.bstart, tl, (const double*)s->gtab, s->ngauss - 1, dinj);
    go(kTTileScan, k_tile_scan, dim3(1), dim3(kTileScanThreads), s->dinst, (const Glob*)s->glob, tl);
    go(kTTile2, f ? k_tile2<true> : k_tile2<false>, dim3(tl.ntile), dim3(kBlock), s->dinst, s->glob, A, B, s->bn,
       tl, s->inv);
  } else {
    go(kTDiffuse, kdiff, dim3(nb), dim3(kBlock), s->dinst, s->glob, A, s->bn.count, (const double*)s->gtab,
       s->ngauss - 1, dinj, s->inv);
    const int ntile = (s->nscan + kHTile - 1) / kHTile;
    go(kTHsum, k_hsum, dim3(ntile), dim3(kHThreads), (const int*)s->bn.count, s->bn.boff, s->nscan);
    const int scanned = ntile > kHFusedTiles;
    if (scanned) go(kTScan2, k_scan2, dim3(1), dim3(kScanBlock), s->bn.boff, ntile);
    go(kTHscan, k_hscan, dim3(ntile), dim3(kHThreads), s->bn.count, (const int*)s->bn.boff, s->bn.bstart, s->nscan,
       scanned);
    go(kTScatter, kscat, dim3(nb), dim3(kBlock), s->dinst, s->glob, A, B, s->bn, s->inv);
  }
  if (s->surf_on)  // first molsort: list changes from adsorption and desorption
    go(kTRelist, k_relist, dim3(s->ninst), dim3(kResolveThreads), s->dinst, s->glob, A, B, s->inv, s->reccap,
       s->bn);
  go(kTPairs, kpair, dim3(nb), dim3(kBlock), s->dinst, s->glob, B, s->bn, s->nscan, s->cand, s->candcap);
  go(kTExact, kexact, dim3(kExactBlocks), dim3(kBlock), s->dinst, s->glob, B, s->bn, s->cand, s->candcap);
  go(kTResolve, kres, dim3(s->ninst), dim3(kResolveThreads), s->dinst, s->glob, A, B, s->inv, s->bn, s->nscan,
     s->bestkey, s->killed, s->reccap);
}

// One step from the current buffer on the stream.  The tile binning needs the
// source buffer in key order; right after an upload or an addition of
// molecules it is not, and the step bins with the counting sort, which leaves
// its destination in order (and the tile queue counters are cleared for the
// tile steps that follow).
static void launch_any(DeviceState* s, const double* dinj, int timed) {
  const bool tile = s->tile_ok && s->sorted;
  launch_step(s, s->cur, dinj, timed, tile);
  if (s->tile_ok && !tile) {
    (void)hipMemsetAsync(s->ta.imm, 0, sizeof(int) * 2 * s->ta.ntile, s->stream);
    (void)hipMemsetAsync(s->ta.stay, 0, sizeof(int) * 2 * s->ta.ntile, s->stream);
    (void)hipMemsetAsync(s->ta.range, 0, sizeof(int2) * 2 * s->ta.ntile, s->stream);
    (void)hipMemsetAsync(s->ta.nspill, 0, sizeof(int) * 2, s->stream);
  }
  s->sorted = true;
}

// Captured step graphs of both buffer parities (with the tile binning when
// the engine has it).
static int capture_graphs(DeviceState* s) {
  for (int k = 0; k < 2; k++) {
    if (s->graph_ready[k]) continue;
    hipGraph_t g;
    HIPOK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    launch_step(s, k, nullptr, 0, s->tile_ok);
    HIPOK(hipStreamEndCapture(s->stream, &g));
    HIPOK(hipGraphInstantiate(&s->graph[k], g, nullptr, nullptr, 0));
    HIPOK(hipGraphDestroy(g));
    s->graph_ready[k] = 1;
  }
  return 0;
}

// Before graph replays: a step that must bin with the counting sort runs
// eagerly; returns the steps it took.
static int sort_step_before_graphs(DeviceState* s, long nsteps) {
  if (nsteps > 0 && s->tile_ok && !s->sorted) {
    launch_any(s, nullptr, 0);
    s->cur ^= 1;
    return 1;
  }
  return 0;
}

// OR of the error bits of all simulations
static int read_errors(DeviceState* s) {
  std::vector<Control> c(s->ninst);
  Glob g;
  if (hipMemcpy(c.data(), s->ctl, sizeof(Control) * s->ninst, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (hipMemcpy(&g, s->glob, sizeof(Glob), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  int e = 0;
  for (auto& x : c) e |= x.err;
  if (g.halt) e |= kErrHalted;
  return e;
}

// ---------------------------------------------------------------------------
// Injected steps that allocate nothing: the deviates are staged in a
// file-scope device array (one per GPU, held until the process ends) instead
// of the engine's injection buffer, which engine_step allocates on first use.
// Calls that overlap on one GPU share the array and race.
constexpr long kNoAllocInjectMax = 3L * 262144;  // doubles: 262144 diffusing molecules, 6 MiB
static __device__ __align__(256) double scratch[kNoAllocInjectMax];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (hipGetSymbolAddress
// once per device; the host-side name is not a device address).
static double* scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch_ptr[dev] = (double*)p;
  }
  return g_scratch_ptr[dev];
}

long engine_noalloc_inject_max() { return kNoAllocInjectMax / 3; }

int engine_step_noalloc(DeviceState* s, const double* inject, long ninject) {
  if (!inject || s->multi) return -1;
  if (ninject * 3 > kNoAllocInjectMax) return -4;
  double* ptr = scratch_ptr();
  if (!ptr) return -1;
  Control c0;
  HIPOK(hipMemcpy(&c0, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  long need = 0;
  for (int ll = 0; ll < s->hinst[0].L.nlist; ll++)
    if (s->htab[0].diffuselist[ll]) need += c0.nl[ll] - c0.nfresh[ll];
  if (ninject < need) return -2;
  if (need > 0) HIPOK(hipMemcpyAsync(ptr, inject, sizeof(double) * 3 * need, hipMemcpyHostToDevice, s->stream));
  launch_any(s, ptr, 0);
  HIPOK(hipGetLastError());
  HIPOK(hipStreamSynchronize(s->stream));
  s->cur ^= 1;
  // error bits of the one simulation, read without host containers
  Control c1;
  Glob g;
  HIPOK(hipMemcpy(&c1, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(&g, s->glob, sizeof(Glob), hipMemcpyDeviceToHost));
  return c1.err | (g.halt ? kErrHalted : 0);
}

int engine_step(DeviceState* s, const double* inject, long ninject) {
  const double* dinj = nullptr;
  if (inject) {
    if (s->multi) return -1;
    Control c0;
    HIPOK(hipMemcpy(&c0, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
    long need = 0;
    for (int ll = 0; ll < s->hinst[0].L.nlist; ll++)
      if (s->htab[0].diffuselist[ll]) need += c0.nl[ll] - c0.nfresh[ll];
    if (ninject < need) return -2;
    if (ninject * 3 > s->injcap) {
      if (s->inj) HIPOK(hipFree(s->inj));
      s->inj = nullptr;
      if (dalloc(&s->inj, (size_t)ninject * 3)) return -1;
      s->injcap = ninject * 3;
    }
    HIPOK(hipMemcpy(s->inj, inject, sizeof(double) * 3 * ninject, hipMemcpyHostToDevice));
    dinj = s->inj;
  }
  launch_any(s, dinj, 0);
  HIPOK(hipGetLastError());
  HIPOK(hipStreamSynchronize(s->stream));
  s->cur ^= 1;
  return read_errors(s);
}

int engine_run(DeviceState* s, long nsteps, int use_graph) {
  hipEvent_t e0 = s->ev[8], e1 = s->ev[9];
  // one captured graph per buffer parity: the launch sequence reads every
  // count from device memory, so the graphs replay for all steps
  if (use_graph && capture_graphs(s)) return -1;
  HIPOK(hipEventRecord(e0, s->stream));
  long k = use_graph ? sort_step_before_graphs(s, nsteps) : 0;
  for (; k < nsteps; k++) {
    if (use_graph) {
      HIPOK(hipGraphLaunch(s->graph[s->cur], s->stream));
      s->cur ^= 1;
    } else {
      launch_any(s, nullptr, 0);
      s->cur ^= 1;
    }
  }
  HIPOK(hipEventRecord(e1, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  float ms = 0;
  HIPOK(hipEventElapsedTime(&ms, e0, e1));
  s->last_ms = ms;
  return read_errors(s);
}

// Species counters of simulation 0 into one observation row:
// [step, count[kMaxSpecies], countf[kMaxSpecies], countk[kMaxSpecies]]
__global__ __launch_bounds__(64) void k_observe(const Control* __restrict__ ctl, int* __restrict__ row) {
  const int t = threadIdx.x;
  if (t == 0) row[0] = (int)ctl->step;
  if (t < kMaxSpecies) {
    row[1 + t] = ctl->count[t];
    row[1 + kMaxSpecies + t] = ctl->countf[t];
    row[1 + 2 * kMaxSpecies + t] = ctl->countk[t];
  }
}

int engine_enqueue(DeviceState* s, long nsteps) {
  if (capture_graphs(s)) return -1;
  for (long k = sort_step_before_graphs(s, nsteps); k < nsteps; k++) {
    HIPOK(hipGraphLaunch(s->graph[s->cur], s->stream));
    s->cur ^= 1;
  }
  return 0;
}

int engine_observe(DeviceState* s, int row) {
  if (row < 0 || row >= kObsRows) return -2;
  if (!s->obs && dalloc(&s->obs, (size_t)kObsRows * kObsStride)) return -1;
  hipLaunchKernelGGL(k_observe, dim3(1), dim3(64), 0, s->stream, s->ctl, s->obs + (size_t)row * kObsStride);
  HIPOK(hipGetLastError());
  return 0;
}

int engine_finish(DeviceState* s) {
  HIPOK(hipStreamSynchronize(s->stream));
  return read_errors(s);
}

int engine_read_obs(DeviceState* s, int nrow, int* out) {
  if (nrow <= 0) return 0;
  if (nrow > kObsRows || !s->obs) return -2;
  HIPOK(hipStreamSynchronize(s->stream));
  HIPOK(hipMemcpy(out, s->obs, sizeof(int) * (size_t)nrow * kObsStride, hipMemcpyDeviceToHost));
  return 0;
}

// Steps with events between the kernels; per-kernel average ms in out[]:
// diffuse, box scan, scatter (and the first molsort with surfaces), event
// overhead (two events with no kernel between them), pair screen, exact
// test, resolve.
int engine_profile(DeviceState* s, long nsteps, double* out) {
  // outputs: diffuse, scan (k_hsum + k_scan2 + k_hscan), scatter, relist,
  // pairs, exact, resolve, tile pass 1, tile offsets, tile pass 2; kernels a
  // step did not launch count 0
  static const int group[kTimedSlots] = {0, 1, 1, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  constexpr int kGroups = 10;
  double acc[kGroups] = {};
  for (long k = 0; k < nsteps; k++) {
    launch_any(s, nullptr, 1);
    HIPOK(hipStreamSynchronize(s->stream));
    s->cur ^= 1;
    for (int j = 0; j < kTimedSlots; j++) {
      if (!(s->timed_mask & (1 << j))) continue;
      float ms = 0;
      HIPOK(hipEventElapsedTime(&ms, s->kev[2 * j], s->kev[2 * j + 1]));
      acc[group[j]] += ms;
    }
  }
  for (int j = 0; j < kGroups; j++) out[j] = acc[j] / (double)(nsteps > 0 ? nsteps : 1);
  return read_errors(s);
}

int engine_get_control(DeviceState* s, int inst, Control* out) {
  if (inst < 0 || inst >= s->ninst) return -2;
  HIPOK(hipStreamSynchronize(s->stream));
  HIPOK(hipMemcpy(out, s->ctl + inst, sizeof(Control), hipMemcpyDeviceToHost));
  return 0;
}

int engine_sync_control(DeviceState* s, Control* out) { return engine_get_control(s, 0, out); }

static int current_nrec(DeviceState* s) {
  Control c;
  Glob g;
  if (hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (hipMemcpy(&g, s->glob, sizeof(Glob), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return g.nrec[(c.step + 1) & 1];
}

int engine_get_list_inst(DeviceState* s, int inst, int ll, int* ident, double* pos, uint64_t* serno, int* box,
                         int max, int* mstate) {
  Control c;
  if (engine_get_control(s, inst, &c)) return -1;
  const Inst& I = s->hinst[inst];
  if (ll < 0 || ll >= I.L.nlist) return -2;
  int n = c.nl[ll];
  if (n > max) n = max;
  if (n <= 0) return c.nl[ll];
  Recs& R = s->buf[s->cur];
  int nb = (s->reccap + kBlock - 1) / kBlock;
  auto kinv = s->multi ? k_rebuild_inv<true> : k_rebuild_inv<false>;
  hipLaunchKernelGGL(kinv, dim3(nb), dim3(kBlock), 0, s->stream, s->dinst, s->glob, R, s->inv);
  HIPOK(hipStreamSynchronize(s->stream));
  std::vector<int> inv(n);
  HIPOK(hipMemcpy(inv.data(), s->inv + I.L.base[ll], sizeof(int) * n, hipMemcpyDeviceToHost));
  int nr = current_nrec(s);
  if (nr <= 0) nr = 1;
  // gather the records of this list only
  std::vector<double> x(nr), y(nr), z(nr);
  std::vector<int> id(nr), bx(nr);
  std::vector<uint64_t> sr(nr);
  HIPOK(hipMemcpy(id.data(), R.ident, sizeof(int) * nr, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(bx.data(), R.box, sizeof(int) * nr, hipMemcpyDeviceToHost));
  if (s->f32) {
    std::vector<float> fx(nr), fy(nr), fz(nr);
    HIPOK(hipMemcpy(fx.data(), R.ox, sizeof(float) * nr, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(fy.data(), R.oy, sizeof(float) * nr, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(fz.data(), R.oz, sizeof(float) * nr, hipMemcpyDeviceToHost));
    for (int r = 0; r < nr; r++) {
      if (bx[r] < 0 || bx[r] >= I.g.nbox) continue;  // not a live record
      double o[3];
      box_origin(I.g, bx[r], o);
      x[r] = o[0] + (double)fx[r];
      y[r] = o[1] + (double)fy[r];
      z[r] = o[2] + (double)fz[r];
    }
  } else {
    HIPOK(hipMemcpy(x.data(), R.px, sizeof(double) * nr, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(y.data(), R.py, sizeof(double) * nr, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(z.data(), R.pz, sizeof(double) * nr, hipMemcpyDeviceToHost));
  }
  HIPOK(hipMemcpy(sr.data(), R.serno, sizeof(uint64_t) * nr, hipMemcpyDeviceToHost));
  std::vector<int> sb;
  if (mstate && R.sb) {
    sb.resize((size_t)nr);
    HIPOK(hipMemcpy(sb.data(), R.sb, sizeof(int) * nr, hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < n; k++) {
    int r = inv[k];
    if (r < 0 || r >= nr) return -3;
    if (mstate) mstate[k] = sb.empty() ? 0 : (sb[(size_t)r] & 3);
    if (ident) ident[k] = id[r];
    if (serno) serno[k] = sr[r];
    if (box) box[k] = bx[r];
    if (pos) {
      pos[3 * k] = x[r];
      pos[3 * k + 1] = y[r];
      pos[3 * k + 2] = z[r];
    }
  }
  return c.nl[ll];
}

int engine_get_list(DeviceState* s, int ll, int* ident, double* pos, uint64_t* serno, int* box, int max) {
  return engine_get_list_inst(s, 0, ll, ident, pos, serno, box, max);
}

int engine_get_log(DeviceState* s, LogRow* rows, int maxrows) {
  Control c;
  if (engine_sync_control(s, &c)) return -1;
  int n = c.nlog < maxrows ? c.nlog : maxrows;
  if (n > 0 && rows) HIPOK(hipMemcpy(rows, s->hinst[0].log, sizeof(LogRow) * n, hipMemcpyDeviceToHost));
  return c.nlog;
}

double engine_last_ms(DeviceState* s) { return s->last_ms; }

int engine_num_instances(DeviceState* s) { return s->ninst; }

}  // namespace smolamd
