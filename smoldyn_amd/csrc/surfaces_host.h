// Host side of the surfaces: the definitions a libsmoldyn caller makes
// (smolAddSurface, smolAddPanel, smolSetSurfaceAction, smolSetSurfaceRate,
// smolSetPanelJump) and their translation into the device tables at
// smolUpdateSim (surfupdateparams, smolsurface.c:3728-3793; box panel lists,
// smolboxes.c:652-671).
#ifndef SMOLDYN_AMD_SURFACES_HOST_H
#define SMOLDYN_AMD_SURFACES_HOST_H

#include <memory>
#include <string>
#include <vector>

#include "engine.h"

namespace smolamd {

constexpr int kSrfMS = 5;   // MSMAX
constexpr int kSrfMS1 = 6;  // MSMAX1

// surfactionstruct (smoldyn.h): per destination state
struct SrfDetail {
  int newspec[kSrfMS1];
  double rate[kSrfMS1], prob[kSrfMS1], cumprob[kSrfMS1], revprob[kSrfMS1];
  int datasrc[kSrfMS1];  // 0 none, 1 rate, 2 probability, 3 species change of a fixed action
};

struct HostPanel {
  std::string name;
  int ps = 0;
  int srf = 0;
  double point[8][3] = {{0}};
  double front[3] = {0, 0, 0};
  int jumpp[2] = {-1, -1};
  int jumpf[2] = {2, 2};
};

struct HostSurface {
  std::string name;
  int action[kMaxSpecies][kSrfMS][3];
  std::unique_ptr<SrfDetail> det[kMaxSpecies][kSrfMS][3];
  std::vector<int> panels[6];  // panel ids per shape, creation order
  int count[6] = {0, 0, 0, 0, 0, 0};
};

struct SurfaceTables {
  std::vector<DevPanel> pnl;
  std::vector<int> bpstart, bplist;
  std::vector<DevSrfAct> act;   // solution-phase molecules, [surface][species][face]
  std::vector<DevSrfAct> actb;  // surface-bound molecules, [surface][species][front/back state]
  int listdiffuse[kMaxList];
  int listbound[kMaxList];
  int blist[kMaxSpecies][2];
};

class SurfaceSet {
 public:
  std::vector<std::unique_ptr<HostSurface>> srf;
  std::vector<HostPanel> pnl;
  double epsilon = 100 * 2.220446049250313e-16;  // surfenablesurfaces, smolsurface.c:1235-1237
  double margin = 100 * 2.220446049250313e-16;
  double neighdist = 10 * 2.220446049250313e-16;

  bool enabled() const { return !srf.empty(); }
  int add_surface(const char* name);
  int find_surface(const char* name) const;
  // ident > 0 one species, -5 all species (1..nspecies-1)
  int set_action(int s, int ident, int nspecies, int ms, int face, int act);
  int set_rate(int s, int ident, int nspecies, int ms, int ms1, int ms2, int newident, double value, int which);
  // returns the panel id, or a negative surfaddpanel error code
  int add_panel(int s, int ps, const char* name, const char* axis, const double* params);
  int find_panel(int s, const char* name) const;
  int set_jump(int p1, int f1, int p2, int f2, int bidirectional);

  // Device tables for a simulation with nspecies species (difc of the
  // solution state), time step dt and the given box grid.  Returns 0, or
  // fills msg and returns nonzero for a model the device path does not run.
  // difc_ms / lookup_ms: diffusion coefficient and list of each species in
  // each molecule state
  int build(const Geometry& g, int nspecies, const double (*difc_ms)[kSrfMS], double dt,
            const int (*lookup_ms)[kSrfMS], const bool* mobile_list, int nlist, SurfaceTables& out, char* msg,
            int msglen) const;
};

}  // namespace smolamd
#endif
