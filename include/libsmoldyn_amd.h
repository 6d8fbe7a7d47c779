/* libsmoldyn_amd: the MI355X (gfx950) engine for Smoldyn's per-timestep
 * particle path behind the libsmoldyn C API.
 *
 * The functions below keep the names, argument meaning, enum values and error
 * behaviour of the reference's libsmoldyn (source/Smoldyn/libsmoldyn.h), so a
 * C/C++ caller that uses this subset links against libsmoldyn_amd.so instead
 * of libsmoldyn.so.  simptr is an opaque handle here: the reference exposes
 * its simstruct, but no caller on the particle path needs its fields (state is
 * read back through smolGetMoleculeCount and the smolamd* inspection calls).
 *
 * Each declaration cites the reference declaration it replaces.
 */
#ifndef LIBSMOLDYN_AMD_H
#define LIBSMOLDYN_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* enum values as source/Smoldyn/libsmoldyn.h:13-26 */
enum ErrorCode {
  ECok = 0,
  ECnotify = -1,
  ECwarning = -2,
  ECnonexist = -3,
  ECall = -4,
  ECmissing = -5,
  ECbounds = -6,
  ECsyntax = -7,
  ECerror = -8,
  ECmemory = -9,
  ECbug = -10,
  ECsame = -11,
  ECwildcard = -12
};

/* source/Smoldyn/smoldyn.h:251-261 */
enum MolecState { MSsoln, MSfront, MSback, MSup, MSdown, MSbsoln, MSall, MSnone, MSsome };

/* source/Smoldyn/smoldyn.h:358-373 */
enum RevParam {
  RPnone, RPirrev, RPconfspread, RPbounce, RPpgem, RPpgemmax, RPpgemmaxw, RPratio, RPunbindrad,
  RPpgem2, RPpgemmax2, RPratio2, RPoffset, RPfixed
};

/* source/Smoldyn/smoldyn.h:484-500 */
enum PanelFace { PFfront, PFback, PFnone, PFboth };
enum PanelShape { PSrect, PStri, PSsph, PScyl, PShemi, PSdisk, PSall, PSnone };

/* source/Smoldyn/smoldyn.h:503-516 */
enum SrfAction { SAreflect, SAtrans, SAabsorb, SAjump, SAport, SAmult, SAno, SAnone, SAadsorb, SArevdes, SAirrevdes, SAflip };

typedef struct smolamd_sim* simptr;
typedef struct smolamd_ensemble* ensembleptr;

/* ---------------- miscellaneous and errors (libsmoldyn.h:34-45) ---------------- */
double smolGetVersion(void);                                                   /* libsmoldyn.h:34 */
void smolSetError(const char* errorfunction, enum ErrorCode errorcode, const char* errorstring,
                  const char* flags);                                          /* libsmoldyn.h:40 */
enum ErrorCode smolGetError(char* errorfunction, char* errorstring, int clearerror); /* libsmoldyn.h:42 */
void smolClearError(void);                                                     /* libsmoldyn.h:43 */
void smolSetDebugMode(int debugmode);                                          /* libsmoldyn.h:44 */
char* smolErrorCodeToString(enum ErrorCode erc, char* string);                 /* libsmoldyn.h:45 */

/* ---------------- sim structure (libsmoldyn.h:49-55) ---------------- */
simptr smolNewSim(int dim, double* lowbounds, double* highbounds);             /* libsmoldyn.h:49 */
enum ErrorCode smolUpdateSim(simptr sim);                                      /* libsmoldyn.h:50 */
enum ErrorCode smolRunTimeStep(simptr sim);                                    /* libsmoldyn.h:51 */
enum ErrorCode smolRunSim(simptr sim);                                         /* libsmoldyn.h:52 */
enum ErrorCode smolRunSimUntil(simptr sim, double breaktime);                  /* libsmoldyn.h:53 */
enum ErrorCode smolFreeSim(simptr sim);                                        /* libsmoldyn.h:54 */

/* ---------------- simulation settings (libsmoldyn.h:65-72) ---------------- */
enum ErrorCode smolSetSimTimes(simptr sim, double timestart, double timestop, double timestep); /* :66 */
enum ErrorCode smolSetTimeStart(simptr sim, double timestart);                 /* libsmoldyn.h:67 */
enum ErrorCode smolSetTimeStop(simptr sim, double timestop);                   /* libsmoldyn.h:68 */
enum ErrorCode smolSetTimeNow(simptr sim, double timenow);                     /* libsmoldyn.h:69 */
enum ErrorCode smolSetTimeStep(simptr sim, double timestep);                   /* libsmoldyn.h:70 */
enum ErrorCode smolSetRandomSeed(simptr sim, long int seed);                   /* libsmoldyn.h:71 */
enum ErrorCode smolSetPartitions(simptr sim, const char* method, double value); /* libsmoldyn.h:72 */

/* ---------------- molecules (libsmoldyn.h:99-118) ---------------- */
enum ErrorCode smolAddSpecies(simptr sim, const char* species, const char* mollist); /* :99 */
int smolGetSpeciesIndex(simptr sim, const char* species);                     /* libsmoldyn.h:100 */
int smolGetSpeciesIndexNT(simptr sim, const char* species);                   /* libsmoldyn.h:101 */
void smolGetSpeciesName(simptr sim, int speciesindex, char* species);         /* libsmoldyn.h:102 */
enum ErrorCode smolSetSpeciesMobility(simptr sim, const char* species, enum MolecState state, double difc,
                                      double* drift, double* difmatrix);      /* libsmoldyn.h:103 */
enum ErrorCode smolAddMolList(simptr sim, const char* mollist);               /* libsmoldyn.h:108 */
int smolGetMolListIndex(simptr sim, const char* mollist);                     /* libsmoldyn.h:109 */
int smolGetMolListIndexNT(simptr sim, const char* mollist);                   /* libsmoldyn.h:110 */
enum ErrorCode smolSetMolList(simptr sim, const char* species, enum MolecState state,
                              const char* mollist);                           /* libsmoldyn.h:112 */
enum ErrorCode smolSetMaxMolecules(simptr sim, int maxmolecules);             /* libsmoldyn.h:113 */
enum ErrorCode smolAddSolutionMolecules(simptr sim, const char* species, int number, double* lowposition,
                                        double* highposition);                /* libsmoldyn.h:114 */
int smolGetMoleculeCount(simptr sim, const char* species, enum MolecState state); /* libsmoldyn.h:117 */

/* ---------------- configuration files (libsmoldyn.h:57-60) ----------------
 * The statements of the models the device path runs (dim, boundaries,
 * low_wall/high_wall, random_seed, time_*, species, difc, molecule_lists,
 * mol_list, molperbox, boxsize, accuracy, max_mol, mol, reaction (incl. <->),
 * reaction_rate, binding_radius, reaction_probability, confspread_radius,
 * product_placement, start_surface ... end_surface with action, rate,
 * rate_internal, panel, jump, output_root, output_files, append_files,
 * output_data, output_precision, output_format, output_file_number, cmd) plus
 * define / ifdefine / ifundefine / else / endif / read_file / end_file;
 * display statements are ignored; anything else fails with file:line in the
 * error message.  Lists are sized when the simulation is updated: a model
 * whose population grows past its initial count plus 25 % (at least 4096)
 * needs max_mol, as the device does not reallocate during a run. */
simptr smolPrepareSimFromFile(const char* filepath, const char* filename, const char* flags); /* :58 */
enum ErrorCode smolLoadSimFromFile(const char* filepath, const char* filename, simptr* simpointer,
                                   const char* flags);                         /* libsmoldyn.h:59 */
enum ErrorCode smolReadConfigString(simptr sim, const char* statement, char* parameters); /* :60 */

/* ---------------- run-time commands (libsmoldyn.h:85-95) ----------------
 * Command timing codes b a @ i x & I j E e N n B A and their scheduling follow
 * the reference (libSteve/SimCommand.c).  The device path runs the commands
 * stop, molcount, molcountheader, molcountspecies, molcountspecieslist,
 * listmols and listmols2; adding any other command returns ECerror.  Observation commands copy the
 * species counters on the device between steps; rows reach files and data
 * tables when the run returns (or every 256 steps), in execution order. */
enum ErrorCode smolSetOutputPath(simptr sim, const char* path);                 /* libsmoldyn.h:87 */
enum ErrorCode smolAddOutputFile(simptr sim, char* filename, int suffix, int append); /* :88 */
enum ErrorCode smolAddOutputData(simptr sim, char* dataname);                   /* libsmoldyn.h:89 */
/* An existing file is not overwritten unless overwrite != 0 (the reference's
   behavior as a Python module; its command-line build asks on the terminal). */
enum ErrorCode smolOpenOutputFiles(simptr sim, int overwrite);                  /* libsmoldyn.h:90 */
enum ErrorCode smolAddCommand(simptr sim, char type, double on, double off, double step, double multiplier,
                              const char* commandstring);                      /* libsmoldyn.h:92 */
enum ErrorCode smolAddCommandFromString(simptr sim, char* string);              /* libsmoldyn.h:93 */
/* *array is allocated with calloc; the caller frees it with free(). */
enum ErrorCode smolGetOutputData(simptr sim, char* dataname, int* nrow, int* ncol, double** array,
                                 int erase);                                   /* libsmoldyn.h:94 */
enum ErrorCode smolRunCommand(simptr sim, const char* commandstring);           /* libsmoldyn.h:95 */

/* ---------------- boundaries and surfaces (libsmoldyn.h:122-136) ---------------- */
enum ErrorCode smolSetBoundaryType(simptr sim, int dimension, int highside, char type); /* :122 */
enum ErrorCode smolAddSurface(simptr sim, const char* surface);               /* libsmoldyn.h:123 */
int smolGetSurfaceIndex(simptr sim, const char* surface);                     /* libsmoldyn.h:124 */
int smolGetSurfaceIndexNT(simptr sim, const char* surface);                   /* libsmoldyn.h:125 */
enum ErrorCode smolSetSurfaceAction(simptr sim, const char* surface, enum PanelFace face, const char* species,
                                    enum MolecState state, enum SrfAction action,
                                    const char* newspecies);                  /* libsmoldyn.h:128 */
enum ErrorCode smolSetSurfaceRate(simptr sim, const char* surface, const char* species, enum MolecState state,
                                  enum MolecState state1, enum MolecState state2, double rate,
                                  const char* newspecies, int isinternal);    /* libsmoldyn.h:129 */
/* Panel shapes on the device path: every 3D shape (PSrect, PStri, PSsph, PScyl,
   PShemi, PSdisk). */
enum ErrorCode smolAddPanel(simptr sim, const char* surface, enum PanelShape panelshape, const char* panel,
                            const char* axisstring, double* params);          /* libsmoldyn.h:130 */
int smolGetPanelIndex(simptr sim, const char* surface, enum PanelShape* panelshapeptr,
                      const char* panel);                                     /* libsmoldyn.h:131 */
int smolGetPanelIndexNT(simptr sim, const char* surface, enum PanelShape* panelshapeptr,
                        const char* panel);                                   /* libsmoldyn.h:132 */
enum ErrorCode smolSetPanelJump(simptr sim, const char* surface, const char* panel1, enum PanelFace face1,
                                const char* panel2, enum PanelFace face2,
                                int isbidirectional);                         /* libsmoldyn.h:134 */
enum ErrorCode smolSetSurfaceSimParams(simptr sim, const char* parameter, double value); /* :136 */

/* ---------------- reactions (libsmoldyn.h:152-158) ---------------- */
enum ErrorCode smolAddReaction(simptr sim, const char* reaction, const char* reactant1, enum MolecState rstate1,
                               const char* reactant2, enum MolecState rstate2, int nproduct,
                               const char** productspecies, enum MolecState* productstates,
                               double rate);                                  /* libsmoldyn.h:152 */
int smolGetReactionIndex(simptr sim, int* orderptr, const char* reaction);    /* libsmoldyn.h:153 */
int smolGetReactionIndexNT(simptr sim, int* orderptr, const char* reaction);  /* libsmoldyn.h:154 */
enum ErrorCode smolSetReactionRate(simptr sim, const char* reaction, double rate, int type); /* :156 */
enum ErrorCode smolSetReactionProducts(simptr sim, const char* reaction, enum RevParam method, double parameter,
                                       const char* product, double* position); /* libsmoldyn.h:158 */

/* ---------------- extensions: no counterpart in the reference ----------------
 * Inspection and parity hooks used by callers that compare the device engine
 * with a CPU path, and the engine's own controls. */

/* The reaction_log statement (smolsim.cpp:2016-2040, RxnSetLog
   smolreact.c:2013-2061; libsmoldyn has no call for it): log the events of
   `reaction` ("all": every reaction) to the output file `filename` when a
   product's serial number is in serno[0..nserno-1] (nserno < 0: every event);
   turnon = 0 removes those serial numbers (all of them for nserno < 0).
   Lines are written by smolRunSim, smolRunSimUntil and smolRunTimeStep. */
enum ErrorCode smolamdSetReactionLog(simptr sim, const char* filename, const char* reaction, const long* serno,
                                     int nserno, int turnon);
/* Initial molecules at given positions (n*dim doubles), serial numbers in order. */
enum ErrorCode smolamdAddMoleculesAt(simptr sim, const char* species, int number, const double* positions);
/* Replace the diffusion draws of the next step by caller-supplied standard
   normal deviates: 3 per diffusing molecule, in the reference's diffuse loop
   order (lists in order, live order within a list).  Like swapping
   sim->diffusefn in the reference (smolsim.cpp:325). */
enum ErrorCode smolamdInjectDisplacements(simptr sim, const double* gauss, long nmol);
/* smolamdInjectDisplacements that allocates no device memory.  No counterpart
   in the reference: it is for callers that cannot allocate (code under HIP
   graph capture, serving loops, harnesses that forbid hipMalloc).  The
   deviates are staged in a fixed device array of 262144 * 3 doubles (6 MiB,
   one per GPU, held until the process ends) instead of the buffer the other
   entry point allocates on first use.  The simulation must already be set up
   by smolUpdateSim (ECerror otherwise: setting it up allocates), and nmol may
   not exceed 262144 diffusing molecules (ECbounds).  Calls that overlap on one
   GPU share the array and race. */
enum ErrorCode smolamdInjectDisplacementsNoAlloc(simptr sim, const double* gauss, long nmol);
/* Record precision of the device path, before smolUpdateSim: 64 (FP64
   positions, bit-exact with the reference's arithmetic) or 32 (each position
   stored as FP32 offsets from the FP64 origin of its virtual box; diffusion,
   walls, surfaces, distance tests and product placement stay FP64).  No
   counterpart in the reference, which stores FP64 positions
   (smoldyn.h:280-293). */
enum ErrorCode smolamdSetPrecision(simptr sim, int bits);
/* Molecules of list `list` in live order; returns the count (or <0). Any of
   the output pointers may be NULL. pos is n*3 doubles. box is the virtual box
   each molecule was assigned to in the last step. */
int smolamdGetMolecules(simptr sim, int list, int* species, double* pos, unsigned long long* serno, int* box,
                        int maxmol);
int smolamdGetListCount(simptr sim);
/* Reaction events of the last step in the order the reference executes them:
   rows of 6 long long: order, reaction index, reactant serials (2), product
   serials (2). Returns the event count. */
int smolamdGetReactionLog(simptr sim, long long* rows, int maxrows);
/* Reaction parameters after smolUpdateSim: out[0..5] = bindrad2, prob,
   unbindrad, prdpos[0][0], prdpos[1][0], rparamt. */
enum ErrorCode smolamdGetReactionParams(simptr sim, int order, int reaction, double* out);
/* Virtual-box grid: side[3], min[3], size[3]; returns the box count. */
int smolamdGetBoxes(simptr sim, int* side, double* boxmin, double* boxsize);
/* The command scheduler on its own, for checking its timing rules: a script
   of lines "times tmin tmax dt", "cmd T on off step multiplier text",
   "cmdstr line" (as smolAddCommandFromString), "dt v", "tmax v", "run"
   (smolRunSim), "runstep" (smolRunTimeStep), "rununtil t" (smolRunSimUntil)
   runs with no simulation attached; a command whose text starts with "stop"
   stops the run.  Each command execution is an event (time, command-list
   position, iteration); each run ends with an event of position
   -1000 - (its return code).  Returns the number of events (or -1). */
int smolamdCommandTimeline(const char* script, int maxev, double* ev_time, int* ev_cmd, long long* ev_iter);
/* Run nsteps time steps (no stopping at time_stop and no run-time commands)
   with the step captured in a HIP graph when use_graph != 0; *ms receives the
   device time (HIP events). */
enum ErrorCode smolamdRunSteps(simptr sim, long nsteps, int use_graph, double* ms);
/* Per-kernel average device time (ms) over nsteps timed steps: each kernel
   is launched with its own start and stop HIP events, stamped at the
   dispatch's start and end.  out[0..9] = diffuse, box scan (histogram sums
   and scan), scatter, first molsort (models with surfaces or absorbing
   walls, else 0), pair screen, exact pair test, resolve, tile binning pass 1,
   tile offsets, tile binning pass 2.  A step bins either with the counting
   sort (diffuse, scan, scatter) or with the tile passes; the other entries
   are 0. */
enum ErrorCode smolamdProfileSteps(simptr sim, long nsteps, double* out);
/* Engine diagnostics: out[0]=live molecules, [1]=edges max, [2]=matching rounds
   of last step, [3]=error bits, [4]=next serial, [5]=step index,
   [6]=molecule-steps run on the device so far, [7]=engine rebuilds for lists
   that filled up (smolRunSim and smolRunTimeStep grow a list that is more
   than half full to twice its capacity, as molexpandlist does,
   smolmolec.c:1911-1950; not for models with surfaces or absorbing walls, nor
   under smolSetMaxMolecules). */
enum ErrorCode smolamdGetDiagnostics(simptr sim, double* out);

/* Ensembles: no counterpart in the reference, which runs replicates as
   separate processes.  n simulations defined with the calls above (not yet
   stepped) run together: every time step is one set of device launches for
   all of them.  Each member gives exactly the results it gives alone.
   Members need the same species and time step and no surfaces; the
   simulations stay owned by the caller and are not stepped themselves. */
ensembleptr smolamdEnsembleNew(simptr* sims, int n);
void smolamdEnsembleFree(ensembleptr e);
int smolamdEnsembleSize(ensembleptr e);
enum ErrorCode smolamdEnsembleRunSteps(ensembleptr e, long nsteps, int use_graph, double* ms);
int smolamdEnsembleGetMoleculeCount(ensembleptr e, int member, const char* species);
int smolamdEnsembleGetMolecules(ensembleptr e, int member, int list, int* species, double* pos,
                                unsigned long long* serno, int* box, int maxmol);
enum ErrorCode smolamdEnsembleGetDiagnostics(ensembleptr e, int member, double* out);

/* Reaction-parameter solver of the host setup (rxnparam.c of the reference). */
double smolamdBindingRadius(double rate, double dt, double difc, double b, int rel);
double smolamdNumRxnRate(double step, double a, double b);
double smolamdUnbindingRadius(double pgem, double dt, double difc, double a);

#ifdef __cplusplus
}
#endif
#endif
