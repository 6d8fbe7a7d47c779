/* CPU oracle (test infrastructure only; never linked into the product).
 * Surface hooks of the oracle simulation.  Simulations without surfaces take
 * the wall path of simulatetimestep (smolsim.cpp:2618-2630); these hooks are
 * the no-surface forms of checksurfaces, checksurfacebound, rxnXsurface and
 * the molecule-on-surface bookkeeping. */
#include <stdlib.h>

#include "osim.h"

void osurf_free(osim* sim) { (void)sim; }
void osurf_removemol(osim* sim, omol* mp, int ll) {
  (void)sim;
  (void)ll;
  mp->pnl = -1;
}
int osurf_addmol(osim* sim, omol* mp, int ll) {
  (void)sim;
  (void)mp;
  (void)ll;
  return 0;
}
void osurf_clearlist(osim* sim, int ll) {
  (void)sim;
  (void)ll;
}
int osurf_update(osim* sim) {
  (void)sim;
  return 0;
}
void osurf_movemol2closepanel(osim* sim, omol* mp) {
  (void)sim;
  (void)mp;
}
int osurf_rxnXsurface(osim* sim, omol* m1, omol* m2) {
  (void)sim;
  (void)m1;
  (void)m2;
  return 0;
}
int osurf_checksurfaces(osim* sim, int ll, int reborn) {
  (void)sim;
  (void)ll;
  (void)reborn;
  return 0;
}
int osurf_checksurfacebound(osim* sim, int ll) {
  (void)sim;
  (void)ll;
  return 0;
}
