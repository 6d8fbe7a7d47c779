/* Checker shim, compiled together with the reference's own libSteve sources
 * into oracle/_ref/libsteve_ref.so (see Makefile).  It only exports wrappers
 * around the reference's header-inline random functions so that tests can
 * read the reference's values; it contains no reference code. */
#include "random2.h"
#include "Zn.h"

void ref_randomize(long seed) { randomize(seed); }
unsigned long ref_randULI(void) { return randULI(); }
double ref_randCOD(void) { return randCOD(); }
double ref_randCCD(void) { return randCCD(); }
void ref_gausstable(double* a, int n, int shuffle) {
  randtableD(a, n, 1);
  if (shuffle) randshuffletableD(a, n);
}
int ref_neighbor7(int* indx, int* c, int* side, int dim, int* wrap, int* mid) {
  int n = neighborZV(indx, c, side, dim, 7, wrap, mid);
  neighborZV(NULL, NULL, NULL, 0, -1, NULL, NULL);
  return n;
}
