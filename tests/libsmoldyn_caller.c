/* A C caller of the engine through include/libsmoldyn_amd.h, written the way
 * the reference's examples/S97_libsmoldyn/test1.c drives libsmoldyn: build a
 * simulation with smolNewSim and the smolSet / smolAdd calls, update it, run
 * it, read molecule counts and output data, free it.  Compiled with gcc by
 * tests/test_capi_host.py (mode "setup": no GPU needed) and
 * tests/test_gpu_commands.py (mode "run").
 *
 * The model: A + B <-> C in a reflective cube of side 40 with a reflecting
 * sphere inside, 2000 A and 2000 B, dt 0.01.  Output, one line each:
 *   step <n> <A> <B> <C>        after each of the first 20 smolRunTimeStep calls
 *   until <A> <B> <C>           after smolRunSimUntil(1.0)
 *   data <rows> <cols>          the molcount table of the run, then its rows
 *   final <A> <B> <C>           after smolRunSim to time 2
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "libsmoldyn_amd.h"

static int fail(const char* what) {
  char fn[256], msg[256];
  enum ErrorCode er = smolGetError(fn, msg, 1);
  printf("error %s: %d %s %s\n", what, (int)er, fn, msg);
  return 1;
}

#define CHECK(call)                     \
  do {                                  \
    if ((call) != ECok) return fail(#call); \
  } while (0)

int main(int argc, char* argv[]) {
  const int run = argc > 1 && !strcmp(argv[1], "run");
  double lo[3] = {0, 0, 0}, hi[3] = {40, 40, 40};
  double sphere[6] = {20, 20, 20, 15, 20, 20};  /* centre, radius, slices, stacks */
  const char* prod[1] = {"C"};
  enum MolecState pst[1] = {MSsoln};
  const char* back[2] = {"A", "B"};
  enum MolecState bst[2] = {MSsoln, MSsoln};
  char fn[256], msg[256];
  int d;

  simptr sim = smolNewSim(3, lo, hi);
  if (!sim) return fail("smolNewSim");
  CHECK(smolSetRandomSeed(sim, 17));
  CHECK(smolSetSimTimes(sim, 0.0, 2.0, 0.01));
  for (d = 0; d < 3; d++) {
    CHECK(smolSetBoundaryType(sim, d, 0, 'r'));
    CHECK(smolSetBoundaryType(sim, d, 1, 'r'));
  }
  CHECK(smolSetPartitions(sim, "molperbox", 4));
  CHECK(smolAddSpecies(sim, "A", ""));
  CHECK(smolAddSpecies(sim, "B", ""));
  CHECK(smolAddSpecies(sim, "C", ""));
  CHECK(smolSetSpeciesMobility(sim, "A", MSsoln, 1.0, NULL, NULL));
  CHECK(smolSetSpeciesMobility(sim, "B", MSsoln, 1.0, NULL, NULL));
  CHECK(smolSetSpeciesMobility(sim, "C", MSsoln, 0.5, NULL, NULL));
  CHECK(smolAddSurface(sim, "ball"));
  CHECK(smolSetSurfaceAction(sim, "ball", PFboth, "all", MSsoln, SAreflect, NULL));
  CHECK(smolAddPanel(sim, "ball", PSsph, NULL, NULL, sphere));
  CHECK(smolAddReaction(sim, "fwd", "A", MSsoln, "B", MSsoln, 1, prod, pst, 20.0));
  CHECK(smolAddReaction(sim, "back", "C", MSsoln, "", MSnone, 2, back, bst, 2.0));
  CHECK(smolSetReactionProducts(sim, "back", RPratio, 1.5, "", NULL));
  CHECK(smolAddSolutionMolecules(sim, "A", 2000, NULL, NULL));
  CHECK(smolAddSolutionMolecules(sim, "B", 2000, NULL, NULL));

  /* the reference's error behaviour: an unknown species is reported by the
     inner lookup and passed on under the caller's name */
  if (smolAddSolutionMolecules(sim, "Q", 5, NULL, NULL) != ECnonexist) return fail("unknown species accepted");
  smolGetError(fn, msg, 1);
  printf("lookup %s | %s\n", fn, msg);
  printf("species %d %d %d\n", smolGetSpeciesIndex(sim, "A"), smolGetSpeciesIndex(sim, "B"),
         smolGetSpeciesIndex(sim, "C"));
  if (!run) {
    printf("setup ok\n");
    CHECK(smolFreeSim(sim));
    return 0;
  }

  CHECK(smolAddOutputData(sim, "counts"));
  CHECK(smolAddCommandFromString(sim, "N 10 molcount counts"));
  CHECK(smolUpdateSim(sim));
  for (d = 1; d <= 20; d++) {
    CHECK(smolRunTimeStep(sim));
    printf("step %d %d %d %d\n", d, smolGetMoleculeCount(sim, "A", MSall), smolGetMoleculeCount(sim, "B", MSall),
           smolGetMoleculeCount(sim, "C", MSall));
  }
  /* smolRunTimeStep leaves the break time half a step ahead (as the
     reference does), so smolRunSim alone would stop after one more step:
     smolRunSimUntil sets the break time and restores it to the stop time,
     after which smolRunSim runs to the end */
  {
    enum ErrorCode er = smolRunSimUntil(sim, 1.0);
    if (er != ECok && er != ECnotify) return fail("smolRunSimUntil");
    printf("until %d %d %d\n", smolGetMoleculeCount(sim, "A", MSall), smolGetMoleculeCount(sim, "B", MSall),
           smolGetMoleculeCount(sim, "C", MSall));
    er = smolRunSim(sim);
    if (er != ECok && er != ECnotify) return fail("smolRunSim");
  }
  {
    int nrow = 0, ncol = 0, r, c;
    double* table = NULL;
    CHECK(smolGetOutputData(sim, "counts", &nrow, &ncol, &table, 0));
    printf("data %d %d\n", nrow, ncol);
    for (r = 0; r < nrow; r++) {
      for (c = 0; c < ncol; c++) printf(c ? " %g" : "%g", table[r * ncol + c]);
      printf("\n");
    }
    free(table);
  }
  printf("final %d %d %d\n", smolGetMoleculeCount(sim, "A", MSall), smolGetMoleculeCount(sim, "B", MSall),
         smolGetMoleculeCount(sim, "C", MSall));
  CHECK(smolFreeSim(sim));
  return 0;
}
