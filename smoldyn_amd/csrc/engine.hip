// HIP kernels of the MI355X Smoldyn particle step (gfx950).
//
// One reference time step (source/Smoldyn/smolsim.cpp:2607-2690
// simulatetimestep, no surfaces) runs as seven launches with no host
// synchronisation:
//
//   k_diffuse   diffuse (smolmolec.c:3115-3168) + checkwalls (smolwall.c:368-424)
//               + pos2box (smolboxes.c:67-79) + box histogram + unireact coins
//               (smolreact.c:2997-3054); one thread per molecule, SoA streams
//   k_scan1/2   exclusive scan of the box histogram (reassignmolecs' box lists)
//   k_scatter   counting-sort scatter into box order
//   k_boxsort   restore live order inside each box (the reference appends in
//               live order, smolboxes.c:908-930) and gather the box-ordered
//               position payload
//   k_pairs     bireact candidate search (smolreact.c:3201-3296): same box and
//               half/full neighbor stencil, FP64 distance tests, emits edges
//               keyed by the reference's loop order
//   k_resolve   one workgroup: greedy matching in the reference's edge order,
//               product creation (doreact, smolreact.c:2662-2901), molsort
//               compaction (smolmolec.c:2943-3023) and product append order
//
// Arithmetic follows the reference operation by operation and the library is
// built with -ffp-contract=off, so FP64 positions match the CPU path bit for
// bit.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

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
constexpr int kScanPer = 4;
constexpr int kScanBlock = 1024;
constexpr int kScanTile = kScanBlock * kScanPer;  // 4096 entries per scan block
constexpr int kResolveThreads = 1024;
constexpr int kLdsSort = 2048;  // entries sorted in LDS by k_resolve

struct Layout {
  int nlist;
  int captot;
  int base[kMaxList + 1];
};

struct Mols {
  double *px, *py, *pz;
  int* ident;
  uint64_t* serno;
  int* box;
  int* rank;
};

struct Binned {
  int* count;   // [nlist*nbox+1], zeroed by k_scan1 after reading
  int* part;    // block-local exclusive prefix
  int* boff;    // exclusive prefix of block sums
  int* bstart;  // flat exclusive prefix: first box-ordered entry of (list, box)
  int* tmp;     // slots in box order, unordered within a box
  int* sslot;   // slots in box order, live order within a box
  int* sbox;
  float4* spk;  // FP32 position + species bits, box order (pair pre-filter)
};

// Product record built per event before the list compaction.
struct Prod {
  int ident;
  int list;
  double x, y, z;
};

struct DeviceState {
  EngineConfig cfg;
  Prod* prods;
  Layout lay;
  Geometry geo;
  DevTables* tab;
  DevTables htab;
  double* gtab;
  int ngauss;
  uint32_t key0, key1;
  Mols m;
  Binned bn;
  int nscan;   // nlist*nbox + 1
  int nscanblk;
  Edge* edges;
  uint64_t* bestkey;
  int* killed;
  int* estate;
  int* accepted;
  UniEvent* uniev;
  LogRow* log;
  uint64_t* sortkey;  // global fallback sort storage
  int* sortval;
  int sortcap;
  int* killslot;
  Control* ctl;
  double* inj;
  long injcap;
  hipStream_t stream;
  hipGraphExec_t graph;
  int graph_ready;
  hipEvent_t ev[16];
  double ktime[8];
  int ktime_n;
  double last_ms;
};

// ---------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ int slot_list(const Layout& L, int t) {
  int ll = 0;
  while (ll + 1 < L.nlist && t >= L.base[ll + 1]) ll++;
  return ll;
}

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

// checkwalls for one molecule, walls in the reference's order 0..5
__device__ __forceinline__ void apply_walls(const Geometry& g, double* p) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int wl = g.wall[2 * d], wh = g.wall[2 * d + 1];
    if (wl == kWallReflect) {
      if (p[d] < g.low[d]) p[d] = 2 * g.low[d] - p[d];
    } else if (wl == kWallPeriodic) {
      if (p[d] < g.low[d]) p[d] += g.high[d] - g.low[d];
    }
    if (wh == kWallReflect) {
      if (p[d] > g.high[d]) p[d] = 2 * g.high[d] - p[d];
    } else if (wh == kWallPeriodic) {
      if (p[d] > g.high[d]) p[d] += g.low[d] - g.high[d];
    }
  }
}

__device__ __forceinline__ uint64_t atomic_load_u64(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int atomic_load_i32(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int gstart(const Binned& bn, int i) { return bn.bstart[i]; }

// ---------------------------------------------------------------------------
// k_diffuse: diffusion, walls, box, histogram, first-order reaction coins

__global__ __launch_bounds__(kBlock) void k_diffuse(Layout L, Geometry g, const DevTables* __restrict__ tab,
                                                    Control* ctl, Mols m, int* __restrict__ bcount,
                                                    const double* __restrict__ gtab, int gmask,
                                                    const double* __restrict__ inj, uint32_t key0, uint32_t key1,
                                                    UniEvent* __restrict__ uniev, int maxevents, int nbox) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.captot) return;
  int ll = slot_list(L, t);
  int p = t - L.base[ll];
  if (p >= ctl->nl[ll]) return;
  int i = m.ident[t];
  uint64_t ser = m.serno[t];
  uint32_t step = ctl->step;
  double pos[3] = {m.px[t], m.py[t], m.pz[t]};
  if (tab->diffuselist[ll]) {
    double ds = tab->difstep[i];
    double gg[3];
    if (inj) {
      long k = (long)ctl->pfx[ll] + p;
      gg[0] = inj[3 * k];
      gg[1] = inj[3 * k + 1];
      gg[2] = inj[3 * k + 2];
    } else {
      smolrng4 r = smolrng_mol(key0, key1, ser, step, SMOLRNG_DIFFUSE, 0);
      gg[0] = gtab[r.v[0] & (uint32_t)gmask];
      gg[1] = gtab[r.v[1] & (uint32_t)gmask];
      gg[2] = gtab[r.v[2] & (uint32_t)gmask];
    }
    pos[0] += ds * gg[0];
    pos[1] += ds * gg[1];
    pos[2] += ds * gg[2];
    apply_walls(g, pos);
    m.px[t] = pos[0];
    m.py[t] = pos[1];
    m.pz[t] = pos[2];
  }
  int b = pos2box(g, pos[0], pos[1], pos[2]);
  m.box[t] = b;
  m.rank[t] = atomicAdd(&bcount[ll * nbox + b], 1);
  // unireact: coin flips in table order, first success reacts
  int nu = tab->nuni[i];
  if (nu > 0) {
    smolrng4 rb;
    int blk = -1;
    for (int j = 0; j < nu; j++) {
      int r = tab->uni[i][j];
      const DevRxn& rx = tab->rxn1[r];
      if (!rx.permit_soln) continue;
      if (blk != j / 4) {
        blk = j / 4;
        rb = smolrng_mol(key0, key1, ser, step, SMOLRNG_UNIRXN, (uint32_t)blk);
      }
      if (smolrng_u01(rb.v[j & 3]) < rx.prob) {
        int e = atomicAdd(&ctl->nuni, 1);
        if (e < maxevents) {
          uniev[e].slot = t;
          uniev[e].r = r;
        } else {
          atomicOr(&ctl->err, kErrEventOverflow);
        }
        m.ident[t] = 0;
        break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// box-list scan: part[i] = exclusive prefix within a 4096-entry tile

__global__ __launch_bounds__(kScanBlock) void k_scan1(int* __restrict__ count, int* __restrict__ part,
                                                      int* __restrict__ bsum, int n) {
  __shared__ int s[kScanBlock];
  int tile0 = blockIdx.x * kScanTile;
  int base = tile0 + threadIdx.x * kScanPer;
  int v[kScanPer];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    int i = base + k;
    v[k] = (i < n) ? count[i] : 0;
    if (i < n) count[i] = 0;
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
  for (int k = 0; k < kScanPer; k++) {
    int i = base + k;
    if (i < n) part[i] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

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

__global__ __launch_bounds__(kBlock) void k_scan3(const int* __restrict__ part, const int* __restrict__ boff,
                                                  int* __restrict__ bstart, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bstart[i] = part[i] + boff[i / kScanTile];
}

__global__ __launch_bounds__(kBlock) void k_scatter(Layout L, const Control* ctl, Mols m, Binned bn, int nbox) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.captot) return;
  int ll = slot_list(L, t);
  if (t - L.base[ll] >= ctl->nl[ll]) return;
  int u = gstart(bn, ll * nbox + m.box[t]) + m.rank[t];
  bn.tmp[u] = t;
}

// Stable order inside each box: rank an entry among its box by slot, which
// is the live-list order of the list segment.
__global__ __launch_bounds__(kBlock) void k_boxsort(Layout L, const Control* ctl, Mols m, Binned bn, int nbox,
                                                    int nscan) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= gstart(bn, nscan - 1)) return;  // live total
  int t = bn.tmp[u];
  int b = m.box[t];
  int ll = slot_list(L, t);
  int i0 = gstart(bn, ll * nbox + b);
  int i1 = gstart(bn, ll * nbox + b + 1);
  int r = 0;
  for (int k = i0; k < i1; k++) r += (bn.tmp[k] < t);
  int dst = i0 + r;
  bn.sslot[dst] = t;
  bn.sbox[dst] = b;
  bn.spk[dst] = make_float4((float)m.px[t], (float)m.py[t], (float)m.pz[t], __int_as_float(m.ident[t]));
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

__device__ __forceinline__ int first_reaction(const DevTables* tab, int i1, int i2, double dist2, uint32_t key0,
                                              uint32_t key1, uint64_t ser1, uint64_t ser2, uint32_t step) {
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
    if (!rx.permit_soln) continue;
    return r;
  }
  return -1;
}

// Exact FP64 test of one candidate (the reference's dist2 in its operation
// order, smolreact.c:3238-3240 and wallcalcdist2 smolwall.c:100-116) and the
// edge for the first reaction that accepts it.
__device__ __forceinline__ void exact_pair(const Geometry& g, const DevTables* tab, Control* ctl, const Mols& m,
                                           Edge* edges, int maxedges, uint32_t key0, uint32_t key1, uint32_t step,
                                           int t1, const double* p1, int i1, uint64_t ser1, int t2, int i2, int wrap,
                                           uint64_t key) {
  double p2[3] = {m.px[t2], m.py[t2], m.pz[t2]};
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
  int r = first_reaction(tab, i1, i2, dist2, key0, key1, ser1, m.serno[t2], step);
  if (r >= 0) emit_edge(ctl, edges, maxedges, key, t1, t2, r, wrap);
}

// One thread per molecule in box order.  Candidates are screened in FP32 from
// one packed 16-byte load each; only pairs within binding radius plus the
// rounding margin reach the exact FP64 test.
__global__ __launch_bounds__(kBlock) void k_pairs(Layout L, Geometry g, const DevTables* __restrict__ tab,
                                                  Control* ctl, Mols m, Binned bn, int nbox, int nscan,
                                                  Edge* __restrict__ edges, int maxedges, uint32_t key0,
                                                  uint32_t key1) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= bn.bstart[nscan - 1]) return;  // live total
  float4 me = bn.spk[u];
  int i1 = __float_as_int(me.w);
  if (i1 == 0) return;
  uint32_t mask = tab->reactmask[i1];
  if (!mask) return;
  int t1 = bn.sslot[u];
  int ll1 = slot_list(L, t1);
  int nl = tab->nlist;
  int b = bn.sbox[u];
  float r2f;
  if (g.prefilter_margin >= 0) {
    double rr = sqrt(tab->maxbind2[i1]) + g.prefilter_margin;
    r2f = (float)(rr * rr);
  } else {
    r2f = 3.0e38f;
  }
  uint64_t ser1 = 0;
  double p1[3];
  bool loaded = false;
  uint32_t step = ctl->step;
  int p1pos = t1 - L.base[ll1];
  int bx = b / (g.side[1] * g.side[2]);
  int by = (b / g.side[2]) % g.side[1];
  int bz = b % g.side[2];
  float wf[3] = {(float)g.width[0], (float)g.width[1], (float)g.width[2]};
  for (int ll2 = ll1; ll2 < nl; ll2++) {
    if (!tab->rxnmol2[ll1 * kMaxList + ll2]) continue;
    uint64_t pairbits = (uint64_t)tab->pairindex[ll1 * kMaxList + ll2] << 59;
    // same box (pass 0): for ll1==ll2 only entries before m1
    {
      int j0 = gstart(bn, ll2 * nbox + b);
      int jend = (ll1 == ll2) ? u : gstart(bn, ll2 * nbox + b + 1);
      for (int k = j0; k < jend; k++) {
        float4 o = bn.spk[k];
        int i2 = __float_as_int(o.w);
        if (!((mask >> i2) & 1u)) continue;
        float dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
        float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > r2f) continue;
        if (!loaded) {
          p1[0] = m.px[t1];
          p1[1] = m.py[t1];
          p1[2] = m.pz[t1];
          ser1 = m.serno[t1];
          loaded = true;
        }
        uint64_t key = pairbits | ((uint64_t)p1pos << 32) | (uint64_t)(k - j0);
        exact_pair(g, tab, ctl, m, edges, maxedges, key0, key1, step, t1, p1, i1, ser1, bn.sslot[k], i2, 0, key);
      }
    }
    // neighbor boxes (pass 1): 3^3 stencil in add2indx3ZV order, half for ll1==ll2
    for (int nb = 0; nb < 27; nb++) {
      if (nb == 13) continue;
      if (ll1 == ll2 && nb > 13) break;
      int c0 = bx + nb / 9 - 1, c1 = by + (nb / 3) % 3 - 1, c2 = bz + nb % 3 - 1;
      int wrap = 0;
      float sh[3] = {0.f, 0.f, 0.f};
      if (c0 < 0) {
        if (!g.wrapflag[0]) continue;
        c0 = g.side[0] - 1;
        wrap += 1;
        sh[0] = -wf[0];
      } else if (c0 >= g.side[0]) {
        if (!g.wrapflag[0]) continue;
        c0 = 0;
        wrap += 2;
        sh[0] = wf[0];
      }
      if (c1 < 0) {
        if (!g.wrapflag[1]) continue;
        c1 = g.side[1] - 1;
        wrap += 1 << 2;
        sh[1] = -wf[1];
      } else if (c1 >= g.side[1]) {
        if (!g.wrapflag[1]) continue;
        c1 = 0;
        wrap += 2 << 2;
        sh[1] = wf[1];
      }
      if (c2 < 0) {
        if (!g.wrapflag[2]) continue;
        c2 = g.side[2] - 1;
        wrap += 1 << 4;
        sh[2] = -wf[2];
      } else if (c2 >= g.side[2]) {
        if (!g.wrapflag[2]) continue;
        c2 = 0;
        wrap += 2 << 4;
        sh[2] = wf[2];
      }
      int b2 = (c0 * g.side[1] + c1) * g.side[2] + c2;
      int j0 = gstart(bn, ll2 * nbox + b2);
      int j1 = gstart(bn, ll2 * nbox + b2 + 1);
      for (int k = j0; k < j1; k++) {
        if (k == u) continue;  // a box that is its own periodic neighbor
        float4 o = bn.spk[k];
        int i2 = __float_as_int(o.w);
        if (!((mask >> i2) & 1u)) continue;
        float dx = (o.x + sh[0]) - me.x, dy = (o.y + sh[1]) - me.y, dz = (o.z + sh[2]) - me.z;
        float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > r2f) continue;
        if (!loaded) {
          p1[0] = m.px[t1];
          p1[1] = m.py[t1];
          p1[2] = m.pz[t1];
          ser1 = m.serno[t1];
          loaded = true;
        }
        uint64_t key = (1ull << 63) | pairbits | ((uint64_t)p1pos << 32) | ((uint64_t)nb << 27) | (uint64_t)(k - j0);
        exact_pair(g, tab, ctl, m, edges, maxedges, key0, key1, step, t1, p1, i1, ser1, bn.sslot[k], i2, wrap, key);
      }
    }
  }
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
          uint64_t a = key[i], c = key[ixj];
          bool up = ((i & k) == 0);
          if ((a > c) == up) {
            key[i] = c;
            key[ixj] = a;
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

__device__ int pow2ceil(int n) {
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

// exclusive scan of one value per thread; returns the prefix, total in *tot
__device__ int wg_scan(int v, int* s, int* tot) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    int x = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  int incl = s[threadIdx.x];
  *tot = s[blockDim.x - 1];
  __syncthreads();
  return incl - v;
}

struct ResolveArgs {
  Layout L;
  Geometry g;
  const DevTables* tab;
  Control* ctl;
  Mols m;
  Edge* edges;
  int maxedges;
  uint64_t* bestkey;
  int* killed;
  int* estate;
  int* accepted;
  UniEvent* uniev;
  int maxevents;
  LogRow* log;
  uint64_t* sk;  // global sort scratch, capacity sortcap
  int* sv;
  int sortcap;
  int* killslot;
  struct Prod* prods;
  uint32_t key0, key1;
};

__global__ __launch_bounds__(kResolveThreads) void k_resolve(ResolveArgs a) {
  __shared__ uint64_t lk[kLdsSort];
  __shared__ int lv[kLdsSort];
  __shared__ int s_scan[kResolveThreads];
  __shared__ int s_flag, s_nacc;
  __shared__ int s_nl[kMaxList], s_nf[kMaxList], s_k0[kMaxList], s_k1[kMaxList], s_pl[kMaxList];
  const int tid = threadIdx.x;
  const int NT = blockDim.x;
  Control* ctl = a.ctl;
  const DevTables* tab = a.tab;
  Mols m = a.m;
  int E = ctl->nedges;
  if (E > a.maxedges) E = a.maxedges;
  int U = ctl->nuni;
  if (U > a.maxevents) U = a.maxevents;
  uint32_t step = ctl->step;

  // ---- greedy matching: accept an edge when it is the smallest live key at
  // both of its molecules; this is the lexicographically-first maximal
  // matching, i.e. what the sequential loop of bireact accepts.
  if (tid == 0) s_nacc = 0;
  for (int e = tid; e < E; e += NT) a.estate[e] = 0;
  __syncthreads();
  int rounds = 0;
  for (;;) {
    if (tid == 0) s_flag = 0;
    __syncthreads();
    for (int e = tid; e < E; e += NT) {
      if (a.estate[e] != 0) continue;
      Edge ed = a.edges[e];
      if (atomic_load_i32(&a.killed[ed.s1]) || atomic_load_i32(&a.killed[ed.s2])) {
        a.estate[e] = 2;
        continue;
      }
      atomicMin((unsigned long long*)&a.bestkey[ed.s1], (unsigned long long)ed.key);
      atomicMin((unsigned long long*)&a.bestkey[ed.s2], (unsigned long long)ed.key);
      s_flag = 1;
    }
    __syncthreads();
    if (!s_flag) break;
    rounds++;
    for (int e = tid; e < E; e += NT) {
      if (a.estate[e] != 0) continue;
      Edge ed = a.edges[e];
      if (atomic_load_u64(&a.bestkey[ed.s1]) == ed.key && atomic_load_u64(&a.bestkey[ed.s2]) == ed.key) {
        a.estate[e] = 1;
        atomicExch(&a.killed[ed.s1], 1);
        atomicExch(&a.killed[ed.s2], 1);
        int k = atomicAdd(&s_nacc, 1);
        a.accepted[k] = e;
      }
    }
    __syncthreads();
    for (int e = tid; e < E; e += NT) {
      if (a.estate[e] == 2) continue;
      Edge ed = a.edges[e];
      atomicExch((unsigned long long*)&a.bestkey[ed.s1], (unsigned long long)kNoKey);
      atomicExch((unsigned long long*)&a.bestkey[ed.s2], (unsigned long long)kNoKey);
    }
    __syncthreads();
  }
  int B = s_nacc;
  __syncthreads();

  // ---- order the events: first-order reactions in live order (slot order),
  // then bimolecular ones in edge-key order
  // the sort regions must fit the scratch arrays: events at [0,U) and
  // [U,U+pow2(B)), kills at [0,pow2(U+2B))
  if (U + pow2ceil(B) > a.sortcap || pow2ceil(U + 2 * B) > a.sortcap) {
    if (tid == 0) atomicOr(&ctl->err, kErrSortOverflow);
    return;
  }
  for (int i = tid; i < U; i += NT) {
    a.sk[i] = (uint64_t)a.uniev[i].slot;
    a.sv[i] = i;
  }
  __syncthreads();
  wg_sort(a.sk, a.sv, U, lk, lv);
  // event order array in sv[0..U) (uni), then accepted edges sorted into sv[U..U+B)
  for (int i = tid; i < B; i += NT) {
    a.sk[U + i] = a.edges[a.accepted[i]].key;
    a.sv[U + i] = a.accepted[i];
  }
  __syncthreads();
  wg_sort(a.sk + U, a.sv + U, B, lk, lv);
  int NE = U + B;

  // ---- product counts and serial numbers in creation order
  // each thread handles a contiguous chunk of events so the scan is in order
  int per = (NE + NT - 1) / NT;
  int e0 = tid * per, e1 = min(NE, e0 + per);
  int myprod = 0;
  for (int e = e0; e < e1; e++) {
    const DevRxn& rx = e < U ? tab->rxn1[a.uniev[a.sv[e]].r] : tab->rxn2[a.edges[a.sv[e]].r];
    myprod += rx.nprod;
  }
  int K;
  int pbase = wg_scan(myprod, s_scan, &K);
  int D = U + 2 * B;  // reactants killed, in molsort's kill count
  uint64_t ser0 = ctl->serno_next;

  // ---- build products (positions, species, lists) and kill reactants
  Prod* prods = a.prods;
  int pc = pbase;
  for (int e = e0; e < e1; e++) {
    double rxnpos[3];
    const DevRxn* rx;
    uint64_t sA, sB = 0;
    int order;
    int ridx;
    if (e < U) {
      UniEvent ev = a.uniev[a.sv[e]];
      rx = &tab->rxn1[ev.r];
      ridx = ev.r;
      order = 1;
      int t = ev.slot;
      rxnpos[0] = m.px[t];
      rxnpos[1] = m.py[t];
      rxnpos[2] = m.pz[t];
      sA = m.serno[t];
      a.killslot[e] = t;
      a.killed[t] = 1;
    } else {
      Edge ed = a.edges[a.sv[e]];
      rx = &tab->rxn2[ed.r];
      ridx = ed.r;
      order = 2;
      double p1[3] = {m.px[ed.s1], m.py[ed.s1], m.pz[ed.s1]};
      double p2[3] = {m.px[ed.s2], m.py[ed.s2], m.pz[ed.s2]};
      int id1 = m.ident[ed.s1], id2 = m.ident[ed.s2];
      if (ed.wrap && !rx->confspread) {  // morebireact: move the faster molecule next to its partner
        double v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          int cc = (ed.wrap >> (2 * d)) & 3;
          if (cc == 0) v[d] = p2[d] - p1[d];
          else if (cc == 1) v[d] = p2[d] - p1[d] - a.g.width[d];
          else v[d] = p2[d] - p1[d] + a.g.width[d];
        }
        if (tab->difc[id1] < tab->difc[id2]) {
#pragma unroll
          for (int d = 0; d < 3; d++) {
            double shift = p1[d] + v[d] - p2[d];
            p2[d] += shift;
          }
        } else {
#pragma unroll
          for (int d = 0; d < 3; d++) {
            double shift = p2[d] - v[d] - p1[d];
            p1[d] += shift;
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
      for (int d = 0; d < 3; d++) rxnpos[d] = x * pA[d] + (1.0 - x) * pB[d];
      sA = m.serno[swap ? ed.s2 : ed.s1];
      sB = m.serno[swap ? ed.s1 : ed.s2];
      a.killslot[U + 2 * (e - U)] = ed.s1;
      a.killslot[U + 2 * (e - U) + 1] = ed.s2;
    }
    double uvec[3] = {0, 0, 0};
    if (rx->rotate) smolrng_unitvec(a.key0, a.key1, sA, step, 3, uvec);
    for (int q = 0; q < rx->nprod; q++) {
      Prod pr;
      pr.ident = rx->prd[q];
      pr.list = rx->prdlist[q];
      double pp[3] = {rxnpos[0], rxnpos[1], rxnpos[2]};
      if (rx->prdpos[q] != 0) {
        double v1[3] = {rx->prdpos[q] * uvec[0], rx->prdpos[q] * uvec[1], rx->prdpos[q] * uvec[2]};
        pp[0] = rxnpos[0] + v1[0];
        pp[1] = rxnpos[1] + v1[1];
        pp[2] = rxnpos[2] + v1[2];
      }
      apply_walls(a.g, pp);  // checkwalls for reborn molecules
      pr.x = pp[0];
      pr.y = pp[1];
      pr.z = pp[2];
      prods[pc] = pr;
      pc++;
    }
    // reaction log
    int li = e;
    LogRow row;
    row.order = order;
    row.r = ridx;
    row.rct[0] = sA;
    row.rct[1] = sB;
    int pfirst = pc - rx->nprod;
    row.prd[0] = rx->nprod > 0 ? ser0 + (uint64_t)pfirst : 0;
    row.prd[1] = rx->nprod > 1 ? ser0 + (uint64_t)pfirst + 1 : 0;
    a.log[li] = row;
  }
  __syncthreads();

  // ---- molsort: sort the killed slots, fill holes from the list tails
  // (swap-from-tail compaction), then append products in the dead-list
  // order the reference produces: creation order reversed and rotated by
  // the number of kills.
  for (int i = tid; i < D; i += NT) {
    a.sk[i] = (uint64_t)a.killslot[i];
    a.sv[i] = 0;
  }
  __syncthreads();
  wg_sort(a.sk, a.sv, D, lk, lv);
  if (tid < a.L.nlist) {
    int ll = tid;
    int lo = a.L.base[ll], hi = a.L.base[ll + 1];
    // range of kills in this list (binary searches on the sorted slots)
    int l = 0, h = D;
    while (l < h) {
      int mid = (l + h) >> 1;
      if ((int)a.sk[mid] < lo) l = mid + 1;
      else h = mid;
    }
    int k0 = l;
    h = D;
    while (l < h) {
      int mid = (l + h) >> 1;
      if ((int)a.sk[mid] < hi) l = mid + 1;
      else h = mid;
    }
    s_k0[ll] = k0;
    s_k1[ll] = l;
    s_nl[ll] = ctl->nl[ll];
    s_nf[ll] = ctl->nl[ll] - (l - k0);
  }
  __syncthreads();
  for (int ll = 0; ll < a.L.nlist; ll++) {
    int k0 = s_k0[ll], k1 = s_k1[ll], nk = k1 - k0;
    int nlc = s_nl[ll], nf = s_nf[ll];
    int lo = a.L.base[ll];
    for (int q = tid; q < nk; q += NT) {
      int t = lo + nf + q;  // tail position nf+q
      // killed?  binary search among this list's kills
      int l = k0, h = k1;
      while (l < h) {
        int mid = (l + h) >> 1;
        if ((int)a.sk[mid] < t) l = mid + 1;
        else h = mid;
      }
      bool dead = (l < k1 && (int)a.sk[l] == t);
      if (dead) continue;
      int above = k1 - l;  // kills at positions > t (t itself not killed)
      int rank = (nlc - 1 - (nf + q)) - above;
      int dst = (int)a.sk[k0 + rank];
      m.px[dst] = m.px[t];
      m.py[dst] = m.py[t];
      m.pz[dst] = m.pz[t];
      m.ident[dst] = m.ident[t];
      m.serno[dst] = m.serno[t];
      m.box[dst] = m.box[t];
    }
  }
  __syncthreads();
  // clear kill marks (slots are reused)
  for (int i = tid; i < D; i += NT) a.killed[a.killslot[i]] = 0;
  // product append order: Q[j] = creation index K-1-((j+D) mod K)
  if (tid < a.L.nlist) s_pl[tid] = 0;
  __syncthreads();
  for (int j0 = 0; j0 < K; j0 += NT) {
    int j = j0 + tid;
    int c = -1, ll = -1;
    if (j < K) {
      c = K - 1 - (int)(((long)j + D) % K);
      ll = prods[c].list;
    }
    for (int l2 = 0; l2 < a.L.nlist; l2++) {
      int tot;
      int pre = wg_scan((ll == l2) ? 1 : 0, s_scan, &tot);
      if (ll == l2) {
        int dst = a.L.base[l2] + s_nf[l2] + s_pl[l2] + pre;
        if (dst < a.L.base[l2 + 1]) {
          Prod pr = prods[c];
          m.px[dst] = pr.x;
          m.py[dst] = pr.y;
          m.pz[dst] = pr.z;
          m.ident[dst] = pr.ident;
          m.serno[dst] = ser0 + (uint64_t)c;
          m.box[dst] = pos2box(a.g, pr.x, pr.y, pr.z);
        } else {
          atomicOr(&ctl->err, kErrCapacity);
        }
      }
      __syncthreads();
      if (tid == 0) s_pl[l2] += tot;
      __syncthreads();
    }
  }
  __syncthreads();
  // ---- species counts, list lengths, counters
  for (int e = tid; e < NE; e += NT) {
    const DevRxn& rx = e < U ? tab->rxn1[a.uniev[a.sv[e]].r] : tab->rxn2[a.edges[a.sv[e]].r];
    atomicSub(&ctl->count[rx.rct0], 1);
    if (e >= U) atomicSub(&ctl->count[rx.rct1], 1);
    for (int q = 0; q < rx.nprod; q++) atomicAdd(&ctl->count[rx.prd[q]], 1);
  }
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    unsigned long long live0 = 0;
    for (int ll = 0; ll < a.L.nlist; ll++) live0 += (unsigned long long)s_nl[ll];
    ctl->molsteps += live0;
    for (int ll = 0; ll < a.L.nlist; ll++) {
      int nlnew = s_nf[ll] + s_pl[ll];
      int cap = a.L.base[ll + 1] - a.L.base[ll];
      if (nlnew > cap) nlnew = cap;
      ctl->nl[ll] = nlnew;
      ctl->pfx[ll] = run;
      if (tab->diffuselist[ll]) run += nlnew;
    }
    ctl->serno_next = ser0 + (uint64_t)K;
    ctl->naccept = B;
    ctl->nlog = NE;
    if (E > ctl->maxedges_seen) ctl->maxedges_seen = E;
    ctl->rounds = rounds;
    ctl->nedges = 0;
    ctl->nuni = 0;
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

static int pow2ceil_h(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

DeviceState* engine_create(const EngineConfig& cfg, const DevTables& tab, const Geometry& geo, const double* gtable,
                           int ngauss, uint32_t key0, uint32_t key1, char* err, int errlen) {
  DeviceState* s = new DeviceState();
  memset((void*)s, 0, sizeof(DeviceState));
  s->cfg = cfg;
  s->geo = geo;
  s->htab = tab;
  s->key0 = key0;
  s->key1 = key1;
  s->ngauss = ngauss;
  s->lay.nlist = cfg.nlist;
  int run = 0;
  for (int ll = 0; ll < cfg.nlist; ll++) {
    s->lay.base[ll] = run;
    run += cfg.cap[ll];
  }
  for (int ll = cfg.nlist; ll <= kMaxList; ll++) s->lay.base[ll] = run;
  s->lay.captot = run;
  int captot = run;
  s->nscan = cfg.nlist * geo.nbox + 1;
  s->nscanblk = (s->nscan + kScanTile - 1) / kScanTile;
  int ok = 0;
  if (s->nscanblk > kScanBlock * 8) {
    snprintf(err, errlen, "too many virtual boxes for the scan (%d)", s->nscan);
    delete s;
    return nullptr;
  }
  s->sortcap = pow2ceil_h(2 * cfg.maxevents + 16);
  ok |= dalloc(&s->m.px, captot) | dalloc(&s->m.py, captot) | dalloc(&s->m.pz, captot);
  ok |= dalloc(&s->m.ident, captot) | dalloc(&s->m.serno, captot) | dalloc(&s->m.box, captot) |
        dalloc(&s->m.rank, captot);
  ok |= dalloc(&s->bn.count, s->nscan) | dalloc(&s->bn.part, s->nscan) | dalloc(&s->bn.boff, s->nscanblk) |
        dalloc(&s->bn.bstart, s->nscan);
  ok |= dalloc(&s->bn.tmp, captot) | dalloc(&s->bn.sslot, captot) | dalloc(&s->bn.sbox, captot) |
        dalloc(&s->bn.spk, captot);
  ok |= dalloc(&s->edges, cfg.maxedges) | dalloc(&s->estate, cfg.maxedges) | dalloc(&s->accepted, cfg.maxedges);
  ok |= dalloc(&s->bestkey, captot) | dalloc(&s->killed, captot);
  ok |= dalloc(&s->uniev, cfg.maxevents) | dalloc(&s->log, s->sortcap);
  ok |= dalloc(&s->sortkey, s->sortcap) | dalloc(&s->sortval, s->sortcap) | dalloc(&s->killslot, s->sortcap);
  ok |= dalloc(&s->ctl, 1) | dalloc(&s->tab, 1) | dalloc(&s->gtab, ngauss);
  ok |= dalloc(&s->prods, (size_t)s->sortcap * kMaxProd);
  if (ok) {
    snprintf(err, errlen, "device allocation failed (%d molecule slots)", captot);
    engine_destroy(s);
    return nullptr;
  }
  s->inj = nullptr;
  s->injcap = 0;
  s->ktime_n = 0;
  (void)hipMemset(s->bn.count, 0, sizeof(int) * s->nscan);
  (void)hipMemset(s->bestkey, 0xFF, sizeof(uint64_t) * captot);
  (void)hipMemset(s->killed, 0, sizeof(int) * captot);
  (void)hipMemset(s->ctl, 0, sizeof(Control));
  (void)hipMemcpy(s->tab, &tab, sizeof(DevTables), hipMemcpyHostToDevice);
  (void)hipMemcpy(s->gtab, gtable, sizeof(double) * ngauss, hipMemcpyHostToDevice);
  (void)hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  for (int k = 0; k < 16; k++) (void)hipEventCreate(&s->ev[k]);
  s->graph_ready = 0;
  return s;
}

void engine_destroy(DeviceState* s) {
  if (!s) return;
  void* ptrs[] = {s->m.px, s->m.py, s->m.pz, s->m.ident, s->m.serno, s->m.box, s->m.rank, s->bn.count, s->bn.part,
                  s->bn.boff, s->bn.bstart, s->bn.tmp, s->bn.sslot, s->bn.sbox, s->bn.spk,
                  s->edges, s->estate, s->accepted, s->bestkey, s->killed, s->uniev, s->log, s->sortkey,
                  s->sortval, s->killslot, s->ctl, s->tab, s->gtab, s->inj, s->prods};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->graph_ready) (void)hipGraphExecDestroy(s->graph);
  if (s->stream) {
    for (int k = 0; k < 16; k++) (void)hipEventDestroy(s->ev[k]);
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
}

int engine_upload_list(DeviceState* s, int ll, int n, const int* ident, const double* pos, const uint64_t* serno) {
  if (ll < 0 || ll >= s->lay.nlist) return -2;
  int cap = s->lay.base[ll + 1] - s->lay.base[ll];
  if (n > cap) return -3;
  int b0 = s->lay.base[ll];
  std::vector<double> x(n), y(n), z(n);
  for (int k = 0; k < n; k++) {
    x[k] = pos[3 * k];
    y[k] = pos[3 * k + 1];
    z[k] = pos[3 * k + 2];
  }
  if (n) {
    HIPOK(hipMemcpy(s->m.px + b0, x.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(s->m.py + b0, y.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(s->m.pz + b0, z.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(s->m.ident + b0, ident, sizeof(int) * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(s->m.serno + b0, serno, sizeof(uint64_t) * n, hipMemcpyHostToDevice));
  }
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  c.nl[ll] = n;
  int run = 0;
  for (int l2 = 0; l2 < s->lay.nlist; l2++) {
    c.pfx[l2] = run;
    if (s->htab.diffuselist[l2]) run += c.nl[l2];
  }
  HIPOK(hipMemcpy(s->ctl, &c, sizeof(Control), hipMemcpyHostToDevice));
  return 0;
}

int engine_set_counters(DeviceState* s, uint64_t serno_next, uint32_t step, const int* counts) {
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  c.serno_next = serno_next;
  c.step = step;
  for (int i = 0; i < kMaxSpecies; i++) c.count[i] = counts ? counts[i] : 0;
  HIPOK(hipMemcpy(s->ctl, &c, sizeof(Control), hipMemcpyHostToDevice));
  return 0;
}

static ResolveArgs resolve_args(DeviceState* s) {
  ResolveArgs a;
  a.L = s->lay;
  a.g = s->geo;
  a.tab = s->tab;
  a.ctl = s->ctl;
  a.m = s->m;
  a.edges = s->edges;
  a.maxedges = s->cfg.maxedges;
  a.bestkey = s->bestkey;
  a.killed = s->killed;
  a.estate = s->estate;
  a.accepted = s->accepted;
  a.uniev = s->uniev;
  a.maxevents = s->cfg.maxevents;
  a.log = s->log;
  a.sk = s->sortkey;
  a.sv = s->sortval;
  a.sortcap = s->sortcap;
  a.killslot = s->killslot;
  a.prods = s->prods;
  a.key0 = s->key0;
  a.key1 = s->key1;
  return a;
}

// Launch one step on s->stream; the number of live molecules for the grids is
// an upper bound (capacity), the kernels read the exact counts on the device.
// timed: record events between the kernels into ev[0..7].
static void launch_step(DeviceState* s, const double* dinj, int timed) {
  int captot = s->lay.captot;
  int nb = (captot + kBlock - 1) / kBlock;
  hipStream_t st = s->stream;
  int nbox = s->geo.nbox;
  if (timed) (void)hipEventRecord(s->ev[0], st);
  hipLaunchKernelGGL(k_diffuse, dim3(nb), dim3(kBlock), 0, st, s->lay, s->geo, s->tab, s->ctl, s->m, s->bn.count,
                     s->gtab, s->ngauss - 1, dinj, s->key0, s->key1, s->uniev, s->cfg.maxevents, nbox);
  if (timed) (void)hipEventRecord(s->ev[1], st);
  hipLaunchKernelGGL(k_scan1, dim3(s->nscanblk), dim3(kScanBlock), 0, st, s->bn.count, s->bn.part, s->bn.boff,
                     s->nscan);
  hipLaunchKernelGGL(k_scan2, dim3(1), dim3(kScanBlock), 0, st, s->bn.boff, s->nscanblk);
  hipLaunchKernelGGL(k_scan3, dim3((s->nscan + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s->bn.part, s->bn.boff,
                     s->bn.bstart, s->nscan);
  if (timed) (void)hipEventRecord(s->ev[2], st);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kBlock), 0, st, s->lay, s->ctl, s->m, s->bn, nbox);
  if (timed) (void)hipEventRecord(s->ev[3], st);
  hipLaunchKernelGGL(k_boxsort, dim3(nb), dim3(kBlock), 0, st, s->lay, s->ctl, s->m, s->bn, nbox, s->nscan);
  if (timed) (void)hipEventRecord(s->ev[4], st);
  hipLaunchKernelGGL(k_pairs, dim3(nb), dim3(kBlock), 0, st, s->lay, s->geo, s->tab, s->ctl, s->m, s->bn, nbox,
                     s->nscan, s->edges, s->cfg.maxedges, s->key0, s->key1);
  if (timed) (void)hipEventRecord(s->ev[5], st);
  hipLaunchKernelGGL(k_resolve, dim3(1), dim3(kResolveThreads), 0, st, resolve_args(s));
  if (timed) (void)hipEventRecord(s->ev[6], st);
}

int engine_step(DeviceState* s, const double* inject, long ninject) {
  const double* dinj = nullptr;
  if (inject) {
    Control c0;
    HIPOK(hipMemcpy(&c0, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
    long need = 0;
    for (int ll = 0; ll < s->lay.nlist; ll++)
      if (s->htab.diffuselist[ll]) need += c0.nl[ll];
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
  launch_step(s, dinj, 0);
  HIPOK(hipGetLastError());
  HIPOK(hipStreamSynchronize(s->stream));
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  return c.err;
}

int engine_run(DeviceState* s, long nsteps, int use_graph) {
  hipEvent_t e0 = s->ev[8], e1 = s->ev[9];
  // The launch sequence reads every count from device memory, so one captured
  // step can be replayed for all steps.
  if (use_graph && !s->graph_ready) {
    hipGraph_t g;
    HIPOK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    launch_step(s, nullptr, 0);
    HIPOK(hipStreamEndCapture(s->stream, &g));
    HIPOK(hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0));
    HIPOK(hipGraphDestroy(g));
    s->graph_ready = 1;
  }
  HIPOK(hipEventRecord(e0, s->stream));
  for (long k = 0; k < nsteps; k++) {
    if (use_graph) HIPOK(hipGraphLaunch(s->graph, s->stream));
    else launch_step(s, nullptr, 0);
  }
  HIPOK(hipEventRecord(e1, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  float ms = 0;
  HIPOK(hipEventElapsedTime(&ms, e0, e1));
  s->last_ms = ms;
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  return c.err;
}

// Run nsteps with events between kernels; per-kernel average ms in out[]:
// diffuse, scan, scatter, boxsort, pairs, resolve.
int engine_profile(DeviceState* s, long nsteps, double* out) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long k = 0; k < nsteps; k++) {
    launch_step(s, nullptr, 1);
    HIPOK(hipStreamSynchronize(s->stream));
    for (int j = 0; j < 6; j++) {
      float ms = 0;
      HIPOK(hipEventElapsedTime(&ms, s->ev[j], s->ev[j + 1]));
      acc[j] += ms;
    }
  }
  for (int j = 0; j < 6; j++) out[j] = acc[j] / (double)(nsteps > 0 ? nsteps : 1);
  Control c;
  HIPOK(hipMemcpy(&c, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  return c.err;
}

int engine_sync_control(DeviceState* s, Control* out) {
  HIPOK(hipStreamSynchronize(s->stream));
  HIPOK(hipMemcpy(out, s->ctl, sizeof(Control), hipMemcpyDeviceToHost));
  return 0;
}

int engine_get_list(DeviceState* s, int ll, int* ident, double* pos, uint64_t* serno, int* box, int max) {
  Control c;
  if (engine_sync_control(s, &c)) return -1;
  if (ll < 0 || ll >= s->lay.nlist) return -2;
  int n = c.nl[ll];
  if (n > max) n = max;
  int b0 = s->lay.base[ll];
  if (n > 0) {
    if (ident) HIPOK(hipMemcpy(ident, s->m.ident + b0, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (serno) HIPOK(hipMemcpy(serno, s->m.serno + b0, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    if (box) HIPOK(hipMemcpy(box, s->m.box + b0, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (pos) {
      std::vector<double> x(n), y(n), z(n);
      HIPOK(hipMemcpy(x.data(), s->m.px + b0, sizeof(double) * n, hipMemcpyDeviceToHost));
      HIPOK(hipMemcpy(y.data(), s->m.py + b0, sizeof(double) * n, hipMemcpyDeviceToHost));
      HIPOK(hipMemcpy(z.data(), s->m.pz + b0, sizeof(double) * n, hipMemcpyDeviceToHost));
      for (int k = 0; k < n; k++) {
        pos[3 * k] = x[k];
        pos[3 * k + 1] = y[k];
        pos[3 * k + 2] = z[k];
      }
    }
  }
  return c.nl[ll];
}

int engine_get_log(DeviceState* s, LogRow* rows, int maxrows) {
  Control c;
  if (engine_sync_control(s, &c)) return -1;
  int n = c.nlog < maxrows ? c.nlog : maxrows;
  if (n > 0 && rows) HIPOK(hipMemcpy(rows, s->log, sizeof(LogRow) * n, hipMemcpyDeviceToHost));
  return c.nlog;
}

double engine_last_ms(DeviceState* s) { return s->last_ms; }

int engine_device_bytes(DeviceState* s, double* out) {
  *out = (double)s->lay.captot * (3 * 8 + 4 + 8 + 4 + 4 + 4 * 4 + 3 * 8 + 8 + 4) + (double)s->nscan * 12;
  return 0;
}

}  // namespace smolamd
