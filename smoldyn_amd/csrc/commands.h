// Run-time commands, output files and output data tables of a simulation.
//
// The scheduler keeps the reference's command superstructure semantics
// (libSteve/SimCommand.c): command templates, one live copy of each in one of
// two queues (float execution times and iteration numbers), the timing codes
// b a @ i x & I j E e N n B A, the queue order (key, then command-list
// position), the re-queue rules of scmdexecute and the end-of-run pass.  The
// simulate loop follows smolsimulate / simulatetimestep (Smoldyn/smolsim.cpp).
//
// Observation commands do not read molecules on the host: an executed
// molcount asks the caller for an observation row (the device copies its
// species counters into a row of a device table between two steps) and
// queues the formatted output; rows are written to files and data tables in
// execution order when the caller flushes.
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <functional>
#include <string>
#include <vector>

namespace smolamd {
namespace cmd {

// enum CMDcode order of libSteve/SimCommand.h:21 (the larger code wins)
enum Code { kOk = 0, kWarn = 1, kPause = 2, kStop = 3, kAbort = 4 };

struct Command {
  char timing = 0;
  double on = 0, off = 0, dt = 0, xt = 0;   // float-time commands
  long long oni = 0, offi = 0, dti = 0;     // iteration commands
  long long invoke = 0;                     // executions so far
  std::string str;                          // command text
};

// Output of one executed observation command, written at flush.
struct Pending {
  int kind = 0;          // 1 molcount, 2 molcountheader, 3 molcountspecies, 4 molcountspecieslist
  double time = 0;
  FILE* fp = nullptr;    // stdout, stderr or an open output file (nullptr: none)
  int data = -1;         // data table, -1 for none
  int row = -1;          // observation row; -1 for none
  uint32_t step = 0;     // device step counter the row must carry
  std::vector<int> sp;   // species of molcountspecies(list)
  std::vector<int> ms;   // their states (0 soln, 1 front, 2 back, 6 all)
};

// Species counters copied from the device: all states, front-bound, back-bound.
struct ObsRow {
  uint32_t step = 0;
  std::vector<int> count, countf, countk;
};

struct Commands {
  int condition = 0;   // 0 new, 1 time step changed, 2 command added, 3 ready
  long long iter = 0;  // executions of the scheduler so far

  struct Template {
    Command c;
    bool live = false;     // a copy is queued
    bool invoked = false;  // a copy was made since the last reset
  };
  std::vector<Template> list;

  struct Entry {
    double kd = 0;
    long long kl = 0;
    int pos = 0;  // command-list position (second key)
    Command c;
  };
  std::vector<Entry> fq, iq;  // float-time and iteration queues, in pop order

  // output files (root and froot prefix each name; suffix _%03i before the extension)
  std::string root, froot;
  struct File {
    std::string name;
    int suffix = 0;
    bool append = false;
    FILE* fp = nullptr;
  };
  std::vector<File> files;
  int precision = -1;    // %g precision, -1 for the default
  char outformat = 's';  // 's' spaces, 'c' commas

  // output data tables (rows of doubles; ncol = widest row)
  struct Data {
    std::string name;
    std::vector<std::vector<double>> rows;
    int ncol = 0;
  };
  std::vector<Data> data;

  std::vector<Pending> pending;

  ~Commands();
};

// ---- set-up (return codes as scmdaddcommand / scmdstr2cmd: 0 ok, 3 syntax,
// 4 not an integer, 5 step not positive, 6 unknown timing)
int add_command(Commands& c, char timing, double on, double off, double step, double mult, const char* str);
int add_command_string(Commands& c, const char* line);
void set_condition(Commands& c, int cond, int upgrade);
// scmdupdatecommands: instantiate and time the commands for [tmin, tmax] at dt
void update(Commands& c, double tmin, double tmax, double dt);

// ---- execution.  Exec runs one command and returns its Code.
using Exec = std::function<int(Command& cmd, double time)>;
// scmdexecute: run every queued command that is due at (time, iteration);
// donow runs every remaining command once.  iter < 0 takes the next iteration.
int execute(Commands& c, double time, double simdt, long long iter, bool donow, const Exec& ex);
// scmdpop: drop float-time commands due at or before t without running them
void pop(Commands& c, double t);

// ---- the simulate loop
struct Clock {
  double tmin = 0, tmax = 100, dt = 0.01, time = 0, tbreak = 1e300;
};
// Step advances the simulation by one time step; a nonzero return ends the
// loop with that code.
using Step = std::function<int()>;
// smolsimulate: commands, then steps until time reaches tmax (1) or tbreak
// (10), a command stops it (7), or Step fails (its code).  Returns the code.
int simulate(Commands& c, Clock& k, const Step& step, const Exec& ex);

// ---- output
// Find a file or data name as scmdgetfptr does for output style 1 (files),
// 2 (data) or 3 (either, plus an optional second word); -1 when not recognized.
int target(Commands& c, const char* words, int outstyle, FILE** fp, int* data);
int open_files(Commands& c, bool overwrite, char* msg, int msglen);
void close_files(Commands& c);
int add_files(Commands& c, const char* names, bool append);
int set_suffix(Commands& c, const char* name, int suffix);
int add_data(Commands& c, const char* names);
int find_data(const Commands& c, const char* name);
// Write the pending outputs in order.  rows[k] is observation row k;
// species names index 1..nspecies-1.  Outputs whose row carries another step
// (stepping stopped early) and all later ones are dropped; returns the number
// written.
int flush(Commands& c, const std::vector<ObsRow>& rows, const std::vector<std::string>& spname);

// One reaction log line (doreact, smolreact.c:2883-2896): time, reaction
// name, reaction position (dim values, %g with the output precision), the
// reactant serial numbers (order of them), "?" for the products before
// product `first`, then the serial numbers of products first..nprod-1.
void write_rxnlog_line(const Commands& c, FILE* fp, double time, const char* rname, const double* pos, int dim,
                       int order, const uint64_t* rct, int first, int nprod, const uint64_t* prd);

// Parse "name" or "name(state)" (molstring2pattern): species index and state
// (0 soln, 1 front, 2 back, 6 all); 0 for "all" species, -1 when not recognized.
int parse_species(const std::vector<std::string>& spname, const char* word, int* ms);

// ---- commands of the device path: stop, molcount, molcountheader,
// molcountspecies, molcountspecieslist, listmols, listmols2
bool supported(const char* str);
// Observe returns the observation row of the current state and the device
// step counter the row will carry, or -1.
using Observe = std::function<int(uint32_t* step)>;
// One molecule of a listmols dump.
struct MolRow {
  int ident = 0, ms = 0;
  double pos[3] = {0, 0, 0};
  uint64_t serno = 0;
};
// Dump writes out every queued output, then returns every molecule in the
// order the reference's molscancmd visits them (lists in order, live order
// within a list); 0 on success.
using Dump = std::function<int(std::vector<MolRow>* mols)>;
// Run one command: a control code, or an observation queued in c.pending
// (listmols output is written at once, after the queued outputs).  kWarn with
// a message in err when the command cannot run.
int run_command(Commands& c, const Command& cmd, double time, const std::vector<std::string>& spname,
                const Observe& observe, const Dump& dump, std::string* err);

}  // namespace cmd
}  // namespace smolamd
