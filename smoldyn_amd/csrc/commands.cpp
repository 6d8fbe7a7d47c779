// Run-time command scheduler, observation commands and output (see commands.h).
#include "commands.h"

#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

namespace smolamd {
namespace cmd {

namespace {

// n-th (1-based) whitespace-separated word of s, nullptr when absent (strnword)
const char* word_at(const char* s, int n) {
  if (!s) return nullptr;
  const char* p = s;
  for (int k = 1;; k++) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    if (!*p) return nullptr;
    if (k == n) return p;
    while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') p++;
  }
}

std::string first_word(const char* s) {
  const char* p = word_at(s, 1);
  if (!p) return std::string();
  const char* q = p;
  while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') q++;
  return std::string(p, q);
}

// position of (key, pos) in a queue kept in pop order: after every entry with
// a smaller key, and after equal keys whose list position is not larger
// (libSteve/queue.c q_insert)
template <typename KeyOf>
size_t insert_at(const std::vector<Commands::Entry>& q, double key, int pos, KeyOf keyof) {
  size_t i = q.size();
  while (i > 0) {
    const Commands::Entry& e = q[i - 1];
    double k2 = keyof(e);
    if (key < k2 || (key == k2 && pos < e.pos)) i--;
    else break;
  }
  return i;
}

void push_float(Commands& c, int pos, const Command& cmd) {
  Commands::Entry e;
  e.kd = cmd.on;
  e.pos = pos;
  e.c = cmd;
  size_t i = insert_at(c.fq, e.kd, pos, [](const Commands::Entry& x) { return x.kd; });
  c.fq.insert(c.fq.begin() + (long)i, e);
}

void push_iter(Commands& c, int pos, const Command& cmd) {
  Commands::Entry e;
  e.kl = cmd.oni;
  e.pos = pos;
  e.c = cmd;
  size_t i = insert_at(c.iq, (double)e.kl, pos, [](const Commands::Entry& x) { return (double)x.kl; });
  c.iq.insert(c.iq.begin() + (long)i, e);
}

bool is_float_timing(char t) { return t && strchr("ba@ix", t); }

// scmddocommandtiming (SimCommand.c:384-420)
void do_timing(Command& cmd, double tmin, double tmax, double dt, long long iter) {
  const long long nsteps = (long long)((tmax - tmin) / dt + 0.5);
  switch (cmd.timing) {
    case 'b':
      cmd.on = cmd.off = tmin - dt;
      cmd.dt = dt;
      break;
    case 'a':
      cmd.on = cmd.off = tmax + dt;
      cmd.dt = dt;
      break;
    case '@':
      cmd.dt = dt;
      break;
    case 'i':
    case 'x':
      if (cmd.on < tmin) cmd.on = tmin;
      if (cmd.off > tmax) cmd.off = tmax;
      break;
    case 'B':
      cmd.oni = cmd.offi = iter - 1;
      cmd.dti = 1;
      break;
    case 'A':
      cmd.oni = cmd.offi = iter + nsteps + 1;
      cmd.dti = 1;
      break;
    case 'I':
    case 'j':
      if (cmd.oni < 0) cmd.oni = iter + 1;
      break;
    case 'E':
    case 'e':
      cmd.oni = iter;
      cmd.offi = iter + nsteps;
      cmd.dti = 1;
      break;
    case 'N':
    case 'n':
      cmd.oni = iter;
      cmd.offi = iter + nsteps;
      break;
    default:
      break;
  }
}

bool is_integer(double v) { return (double)(long long)v == v; }

// plain number word; false when the word is not a number
bool read_number(const char*& p, double* v) {
  while (*p == ' ' || *p == '\t') p++;
  char* end = nullptr;
  errno = 0;
  double x = strtod(p, &end);
  if (end == p || (*end && *end != ' ' && *end != '\t' && *end != '\n' && *end != '\r')) return false;
  *v = x;
  p = end;
  return true;
}

// "%g" with the output precision, separators by the output format
void write_num(const Commands& c, FILE* fp, double v) {
  if (!fp) return;
  if (c.precision >= 0) fprintf(fp, "%.*g", c.precision, v);
  else fprintf(fp, "%g", v);
}
void write_sep(const Commands& c, FILE* fp) {
  if (fp) fputc(c.outformat == 'c' ? ',' : ' ', fp);
}

void data_append(Commands& c, int did, bool newrow, double v) {
  if (did < 0) return;
  Commands::Data& d = c.data[(size_t)did];
  if (newrow || d.rows.empty()) d.rows.emplace_back();
  d.rows.back().push_back(v);
  if ((int)d.rows.back().size() > d.ncol) d.ncol = (int)d.rows.back().size();
}

int find_name(const std::vector<std::string>& names, const std::string& s) {
  for (size_t i = 0; i < names.size(); i++)
    if (names[i] == s) return (int)i;
  return -1;
}

int find_file(const Commands& c, const std::string& s) {
  for (size_t i = 0; i < c.files.size(); i++)
    if (c.files[i].name == s) return (int)i;
  return -1;
}

std::string file_path(const Commands& c, const Commands::File& f) {
  // root + froot + name, with _%03i before the extension when the suffix is set
  std::string s = c.root + c.froot;
  size_t dot = f.name.rfind('.');
  std::string stem = dot == std::string::npos ? f.name : f.name.substr(0, dot);
  s += stem;
  if (f.suffix) {
    char b[16];
    snprintf(b, sizeof(b), "_%03i", f.suffix);
    s += b;
  }
  if (dot != std::string::npos) s += f.name.substr(dot);
  return s;
}

}  // namespace

Commands::~Commands() { close_files(*this); }

void set_condition(Commands& c, int cond, int upgrade) {
  if (upgrade == 0 && c.condition > cond) c.condition = cond;
  else if (upgrade == 1 && c.condition < cond) c.condition = cond;
  else if (upgrade == 2) c.condition = cond;
}

// scmdaddcommand (SimCommand.c:273-336)
int add_command(Commands& c, char timing, double on, double off, double step, double mult, const char* str) {
  if (!str || !str[0]) return 3;
  Command cmd;
  cmd.timing = timing;
  if (timing && strchr("baBAEe", timing)) {
  } else if (timing == '@') {
    cmd.on = cmd.off = on;
  } else if (timing == 'i' || timing == 'x') {
    if (step <= 0) return 5;
    cmd.on = on;
    cmd.off = off;
    cmd.dt = step;
    if (timing == 'x') cmd.xt = mult;
  } else if (timing == '&') {
    if (!is_integer(on)) return 4;
    cmd.oni = cmd.offi = (long long)on;
    cmd.dti = 1;
  } else if (timing == 'I' || timing == 'j') {
    if (!is_integer(on) || !is_integer(off) || !is_integer(step)) return 4;
    if (step <= 0) return 5;
    cmd.oni = (long long)on;
    cmd.offi = (long long)off;
    cmd.dti = (long long)step;
  } else if (timing == 'N' || timing == 'n') {
    if (!is_integer(step)) return 4;
    if (step <= 0) return 5;  // a zero period would repeat within one iteration without end
    cmd.dti = (long long)step;
  } else {
    return 6;
  }
  cmd.str = str;
  if (!cmd.str.empty() && cmd.str.back() == '\n') cmd.str.pop_back();
  Commands::Template t;
  t.c = cmd;
  c.list.push_back(t);
  set_condition(c, 2, 0);
  return 0;
}

// scmdstr2cmd (SimCommand.c:340-380): "timing [on off step [multiplier]] command"
int add_command_string(Commands& c, const char* line) {
  if (!line) return 3;
  const char* p = line;
  while (*p == ' ' || *p == '\t') p++;
  if (!*p) return 3;
  char timing = *p;
  const char* rest = word_at(p, 2);
  if (!rest) return 3;
  double v[4] = {0, 0, 0, 0};
  int nv;
  if (strchr("baBAEe", timing)) nv = 0;
  else if (timing == '@' || timing == '&') nv = 1;
  else if (timing == 'N' || timing == 'n') nv = -1;  // one number: the step
  else if (timing == 'i' || timing == 'I' || timing == 'j') nv = 3;
  else if (timing == 'x') nv = 4;
  else return 6;
  const char* q = rest;
  if (nv == -1) {
    if (!read_number(q, &v[2])) return 3;
  } else {
    for (int k = 0; k < nv; k++)
      if (!read_number(q, &v[k])) return 3;
  }
  if (nv != 0) {
    q = word_at(q, 1);
    if (!q) return 3;
  }
  return add_command(c, timing, v[0], v[1], v[2], v[3], q);
}

// scmdupdatecommands (SimCommand.c:423-477)
void update(Commands& c, double tmin, double tmax, double dt) {
  if (c.condition == 3) return;
  if (dt <= 0 || tmax < tmin) return;
  if (c.condition == 0) {
    c.fq.clear();
    c.iq.clear();
    for (auto& t : c.list) t.live = t.invoked = false;
  }
  for (size_t i = 0; i < c.list.size(); i++) {
    Commands::Template& t = c.list[i];
    if (!t.live && !t.invoked) {
      Command cmd = t.c;
      do_timing(cmd, tmin, tmax, dt, c.iter);
      if (is_float_timing(cmd.timing)) push_float(c, (int)i, cmd);
      else push_iter(c, (int)i, cmd);
      t.live = t.invoked = true;
    } else if (c.condition == 1 && t.live) {
      // a changed time step re-times the queued copy; its queue key stays
      for (auto* q : {&c.fq, &c.iq})
        for (auto& e : *q)
          if (e.pos == (int)i) do_timing(e.c, tmin, tmax, dt, c.iter);
    }
  }
  set_condition(c, 3, 1);
}

// scmdpop (SimCommand.c:480-491)
void pop(Commands& c, double t) {
  while (!c.fq.empty() && c.fq.front().kd <= t) {
    c.list[(size_t)c.fq.front().pos].live = false;
    c.fq.erase(c.fq.begin());
  }
}

// scmdexecute (SimCommand.c:494-545)
int execute(Commands& c, double time, double simdt, long long iter, bool donow, const Exec& ex) {
  int code2 = kOk;
  if (iter < 0) iter = c.iter++;
  else c.iter = iter;
  while (!c.iq.empty() && (c.iq.front().kl <= iter || donow)) {
    Commands::Entry e = c.iq.front();
    c.iq.erase(c.iq.begin());
    e.c.invoke++;
    int code1 = ex(e.c, time);
    if (e.c.oni + e.c.dti <= e.c.offi && !donow && (code1 == kOk || code1 == kPause)) {
      e.c.oni += e.c.dti;
      push_iter(c, e.pos, e.c);
    } else {
      c.list[(size_t)e.pos].live = false;
    }
    if (code1 == kAbort) return code1;
    if (code1 > code2) code2 = code1;
  }
  while (!c.fq.empty() && (c.fq.front().kd <= time || donow)) {
    Commands::Entry e = c.fq.front();
    c.fq.erase(c.fq.begin());
    e.c.invoke++;
    int code1 = ex(e.c, time);
    double dt = e.c.dt >= simdt ? e.c.dt : simdt;
    if (e.c.on + dt <= e.c.off && !donow && (code1 == kOk || code1 == kPause)) {
      e.c.on += dt;
      if (e.c.xt > 1) e.c.dt *= e.c.xt;
      push_float(c, e.pos, e.c);
    } else {
      c.list[(size_t)e.pos].live = false;
    }
    if (code1 == kAbort) return code1;
    if (code1 > code2) code2 = code1;
  }
  return code2;
}

// smolsimulate (smolsim.cpp:2747-2760) with the time handling of
// simulatetimestep (:2680-2690) and simdocommands (:2572-2582)
int simulate(Commands& c, Clock& k, const Step& step, const Exec& ex) {
  int er = 0;
  int code = execute(c, k.time, k.dt, -1, false, ex);
  if (code == kStop || code == kAbort) er = 7;
  while (!er) {
    int s = step();
    if (s) {
      er = s;
      break;
    }
    k.time += k.dt;
    code = execute(c, k.time, k.dt, -1, false, ex);
    if (code == kStop || code == kAbort) er = 7;
    else if (k.time >= k.tmax) er = 1;
    else if (k.time >= k.tbreak) er = 10;
  }
  if (er != 10) {
    pop(c, k.tmax);
    execute(c, k.time, k.dt, -1, true, ex);
    set_condition(c, 0, 0);
  }
  return er;
}

// scmdgetfptr (SimCommand.c:949-988)
int target(Commands& c, const char* words, int outstyle, FILE** fp, int* data) {
  if (fp) *fp = nullptr;
  if (data) *data = -1;
  if (outstyle == 0) return 0;
  std::vector<std::string> dn;
  for (auto& d : c.data) dn.push_back(d.name);
  int fid = -1, did = -1;
  bool std_stream = false;
  std::string name = first_word(words);
  if (name.empty()) {
    *fp = stdout;
    std_stream = true;
  } else if (fp && name == "stdout") {
    *fp = stdout;
    std_stream = true;
  } else if (fp && name == "stderr") {
    *fp = stderr;
    std_stream = true;
  } else {
    if (fp) fid = find_file(c, name);
    if (data) did = find_name(dn, name);
  }
  if (fp && std_stream && (outstyle & 1)) {
  } else if (fid >= 0 && (outstyle & 1)) {
    *fp = c.files[(size_t)fid].fp;
  } else if (did >= 0 && (outstyle & 2)) {
    *data = did;
  } else {
    return -1;
  }
  if (outstyle < 3) return 1;
  const char* w2 = word_at(words, 2);
  if (!w2) return 1;
  fid = did = -1;
  name = first_word(w2);
  if (fp && !*fp && name == "stdout") *fp = stdout;
  else if (fp && !*fp && name == "stderr") *fp = stderr;
  else if (fp && !*fp) fid = find_file(c, name);
  else if (data && *data < 0) did = find_name(dn, name);
  if (fid >= 0) *fp = c.files[(size_t)fid].fp;
  else if (did >= 0) *data = did;
  else return 1;
  return 2;
}

// scmdopenfiles (SimCommand.c:862-904); an existing file is never overwritten
// without the overwrite flag (the reference's Python-module behavior)
int open_files(Commands& c, bool overwrite, char* msg, int msglen) {
  close_files(c);
  for (auto& f : c.files) {
    if (f.name == "stdout") {
      f.fp = stdout;
      continue;
    }
    if (f.name == "stderr") {
      f.fp = stderr;
      continue;
    }
    std::string path = file_path(c, f);
    if (!f.append && !overwrite) {
      FILE* t = fopen(path.c_str(), "r");
      if (t) {
        fclose(t);
        if (msg) snprintf(msg, (size_t)msglen, "file '%s' already exists; refusing to overwrite", f.name.c_str());
        return 1;
      }
    }
    f.fp = fopen(path.c_str(), f.append ? "a" : "w");
    if (!f.fp) {
      if (msg) snprintf(msg, (size_t)msglen, "failed to open file '%s' for writing (errno %d)", path.c_str(), errno);
      return 1;
    }
  }
  return 0;
}

void close_files(Commands& c) {
  for (auto& f : c.files) {
    if (f.fp && f.fp != stdout && f.fp != stderr) fclose(f.fp);
    f.fp = nullptr;
  }
}

// scmdsetfnames (SimCommand.c:787-848)
int add_files(Commands& c, const char* names, bool append) {
  for (int k = 1;; k++) {
    const char* w = word_at(names, k);
    if (!w) break;
    Commands::File f;
    f.name = first_word(w);
    f.append = append;
    c.files.push_back(f);
  }
  return 0;
}

int set_suffix(Commands& c, const char* name, int suffix) {
  int fid = find_file(c, name ? first_word(name) : std::string());
  if (fid < 0) return 1;
  c.files[(size_t)fid].suffix = suffix;
  return 0;
}

// scmdsetdnames (SimCommand.c:717-761): a repeated name gets another table
int add_data(Commands& c, const char* names) {
  for (int k = 1;; k++) {
    const char* w = word_at(names, k);
    if (!w) break;
    Commands::Data d;
    d.name = first_word(w);
    c.data.push_back(d);
  }
  return 0;
}

int find_data(const Commands& c, const char* name) {
  if (!name) return -1;
  for (size_t i = 0; i < c.data.size(); i++)
    if (c.data[i].name == name) return (int)i;
  return -1;
}

int parse_species(const std::vector<std::string>& spname, const char* word, int* ms) {
  std::string w = first_word(word);
  *ms = 0;
  if (w.empty()) return -1;
  if (w.back() == ')') {
    size_t o = w.rfind('(');
    if (o == std::string::npos) return -1;
    std::string st = w.substr(o + 1, w.size() - o - 2);
    w = w.substr(0, o);
    if (st == "soln" || st == "solution") *ms = 0;
    else if (st == "front") *ms = 1;
    else if (st == "back") *ms = 2;
    else if (st == "all") *ms = 6;
    else return -1;
  }
  if (w == "all") return 0;  // every species
  for (size_t i = 1; i < spname.size(); i++)
    if (spname[i] == w) return (int)i;
  return -1;
}

bool supported(const char* str) {
  std::string w = first_word(str);
  return w.empty() || w == "stop" || w == "molcount" || w == "molcountheader" || w == "molcountspecies" ||
         w == "molcountspecieslist" || w == "listmols" || w == "listmols2";
}

namespace {

// molserno2string (smolmolec.c:625-630)
void write_serno(FILE* fp, uint64_t s) {
  if (!fp) return;
  if (s < 0xFFFFFFFFull) fprintf(fp, "%llu", (unsigned long long)s);
  else fprintf(fp, "%llu.%llu", (unsigned long long)(s >> 32), (unsigned long long)(s & 0xFFFFFFFFull));
}

const char* ms_name(int ms) {  // molms2string (smolmolec.c:93-103)
  static const char* names[] = {"solution", "front", "back", "up", "down", "bsoln", "all", "none", "some"};
  return ms >= 0 && ms <= 8 ? names[ms] : "none";
}

}  // namespace

// cmdlistmols (smolcmd.c:2044-2075): "name(state) x y z serno", data rows
// ident, state, x, y, z, serno.  cmdlistmols2 (:2079-2110): the command's
// execution count, ident and state as numbers first.
static int listmols(Commands& c, const Command& cmd, bool two, const char* line2,
                    const std::vector<std::string>& spname, const Dump& dump, std::string* err) {
  FILE* fp = nullptr;
  int did = -1;
  if (target(c, line2, 3, &fp, &did) == -1) {
    if (err) *err = "file or data name not recognized";
    return kWarn;
  }
  std::vector<MolRow> mols;
  if (dump(&mols)) {
    if (err) *err = "reading the molecules from the device failed";
    return kWarn;
  }
  for (const MolRow& m : mols) {
    if (m.ident <= 0) continue;
    if (two) {
      if (fp) fprintf(fp, "%lld", (long long)cmd.invoke);
      write_sep(c, fp);
      if (fp) fprintf(fp, "%i", m.ident);
      write_sep(c, fp);
      if (fp) fprintf(fp, "%i", m.ms);
      data_append(c, did, true, (double)cmd.invoke);
      data_append(c, did, false, (double)m.ident);
      data_append(c, did, false, (double)m.ms);
    } else {
      if (fp) fprintf(fp, "%s(%s)", m.ident < (int)spname.size() ? spname[(size_t)m.ident].c_str() : "?",
                      ms_name(m.ms));
      data_append(c, did, true, (double)m.ident);
      data_append(c, did, false, (double)m.ms);
    }
    for (int d = 0; d < 3; d++) {
      write_sep(c, fp);
      write_num(c, fp, m.pos[d]);
      data_append(c, did, false, m.pos[d]);
    }
    write_sep(c, fp);
    write_serno(fp, m.serno);
    data_append(c, did, false, (double)m.serno);
    if (fp) fputc('\n', fp);
  }
  if (fp) fflush(fp);
  return kOk;
}

// docommand (smolcmd.c:175-320) for the commands of the device path
int run_command(Commands& c, const Command& cmd, double time, const std::vector<std::string>& spname,
                const Observe& observe, const Dump& dump, std::string* err) {
  const char* line = cmd.str.c_str();
  std::string w = first_word(line);
  if (w.empty()) return kOk;
  const char* line2 = word_at(line, 2);
  if (w == "stop") return kStop;  // cmdstop (smolcmd.c:326-330)
  if (w == "listmols" || w == "listmols2") return listmols(c, cmd, w == "listmols2", line2, spname, dump, err);
  Pending p;
  p.time = time;
  auto fail = [&](const char* m) {
    if (err) *err = m;
    return (int)kWarn;
  };
  if (w == "molcount") {
    if (target(c, line2, 3, &p.fp, &p.data) == -1) return fail("file or data name not recognized");
    p.kind = 1;
  } else if (w == "molcountheader") {
    if (target(c, line2, 1, &p.fp, nullptr) == -1) return fail("file name not recognized");
    p.kind = 2;
  } else if (w == "molcountspecies") {
    int ms = 0;
    const int i = parse_species(spname, line2, &ms);
    if (i < 0) return fail("species is missing or cannot be read");
    if (target(c, word_at(line2, 2), 3, &p.fp, &p.data) == -1) return fail("file or data name not recognized");
    p.kind = 3;
    p.sp.push_back(i);
    p.ms.push_back(ms);
  } else if (w == "molcountspecieslist") {
    if (target(c, line2, 3, &p.fp, &p.data) == -1) return fail("file or data name not recognized");
    for (int k = 2; word_at(line2, k); k++) {
      int ms = 0;
      const int i = parse_species(spname, word_at(line2, k), &ms);
      if (i < 0) return fail("species is missing or cannot be read");
      p.sp.push_back(i);
      p.ms.push_back(ms);
    }
    p.kind = 4;
  } else {
    return fail("command not supported by the device engine");
  }
  if (p.kind != 2) {
    p.row = observe(&p.step);
    if (p.row < 0) return fail("observation of the species counters failed");
  }
  c.pending.push_back(p);
  return kOk;
}

int flush(Commands& c, const std::vector<ObsRow>& rows, const std::vector<std::string>& spname) {
  int written = 0;
  auto count_of = [](const ObsRow& r, int i, int ms) {
    const int all = r.count[(size_t)i], f = r.countf[(size_t)i], k = r.countk[(size_t)i];
    return ms == 6 ? all : ms == 1 ? f : ms == 2 ? k : all - f - k;
  };
  auto count_sp = [&](const ObsRow& r, int i, int ms) {
    if (i > 0) return count_of(r, i, ms);
    int s = 0;
    for (size_t j = 1; j < spname.size(); j++) s += count_of(r, (int)j, ms);
    return s;
  };
  for (auto& p : c.pending) {
    const ObsRow* r = nullptr;
    if (p.row >= 0) {
      if (p.row >= (int)rows.size() || rows[(size_t)p.row].step != p.step) break;
      r = &rows[(size_t)p.row];
    }
    switch (p.kind) {
      case 1: {  // cmdmolcount (smolcmd.c:848-904): time, then each species, all states
        write_num(c, p.fp, p.time);
        data_append(c, p.data, true, p.time);
        for (size_t i = 1; i < spname.size(); i++) {
          write_sep(c, p.fp);
          int n = r->count[i];
          if (p.fp) fprintf(p.fp, "%i", n);
          data_append(c, p.data, false, (double)n);
        }
        if (p.fp) fputc('\n', p.fp);
        break;
      }
      case 2: {  // cmdmolcountheader (smolcmd.c:831-844)
        if (p.fp) {
          fputs("time", p.fp);
          for (size_t i = 1; i < spname.size(); i++) {
            write_sep(c, p.fp);
            fputs(spname[i].c_str(), p.fp);
          }
          fputc('\n', p.fp);
        }
        break;
      }
      case 3:    // cmdmolcountspecies (smolcmd.c:1972-1996)
      case 4: {  // cmdmolcountspecieslist (smolcmd.c:1999-2026)
        write_num(c, p.fp, p.time);
        data_append(c, p.data, true, p.time);
        for (size_t q = 0; q < p.sp.size(); q++) {
          write_sep(c, p.fp);
          int n = count_sp(*r, p.sp[q], p.ms[q]);
          if (p.fp) fprintf(p.fp, "%i", n);
          data_append(c, p.data, false, (double)n);
        }
        if (p.fp) fputc('\n', p.fp);
        break;
      }
      default:
        break;
    }
    if (p.fp) fflush(p.fp);
    written++;
  }
  c.pending.clear();
  return written;
}

void write_rxnlog_line(const Commands& c, FILE* fp, double time, const char* rname, const double* pos, int dim,
                       int order, const uint64_t* rct, int first, int nprod, const uint64_t* prd) {
  if (!fp) return;
  write_num(c, fp, time);
  fprintf(fp, " %s", rname);
  for (int d = 0; d < dim; d++) {
    fputc(' ', fp);
    write_num(c, fp, pos[d]);
  }
  for (int k = 0; k < order; k++) {
    fputc(' ', fp);
    write_serno(fp, rct[k]);
  }
  for (int q = 0; q < first; q++) fprintf(fp, " ?");
  for (int q = first; q < nprod; q++) {
    fputc(' ', fp);
    write_serno(fp, prd[q]);
  }
  fputc('\n', fp);
}

}  // namespace cmd
}  // namespace smolamd
