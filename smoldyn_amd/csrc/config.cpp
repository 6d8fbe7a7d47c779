// Configuration files on the device path: the reference's loadsim and
// simreadstring (Smoldyn/smolsim.cpp:626-2263, surfreadstring in
// smolsurface.c) for the statements of the models this engine runs.  Each
// statement goes through the same libsmoldyn calls a C caller would make, in
// file order (molecule placement follows random_seed as in the reference).
// Display statements are accepted and ignored; any other statement is an
// error that names it and its line.
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "config_internal.h"

namespace smolamd {

namespace {

std::vector<std::string> split(const std::string& s) {
  std::vector<std::string> w;
  std::istringstream is(s);
  std::string t;
  while (is >> t) w.push_back(t);
  return w;
}

bool num(const std::string& t, double* v) {
  if (t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  double x = strtod(t.c_str(), &end);
  if (!end || *end || errno == ERANGE) return false;
  *v = x;
  return true;
}

bool inum(const std::string& t, long* v) {
  if (t.empty()) return false;
  char* end = nullptr;
  long x = strtol(t.c_str(), &end, 10);
  if (!end || *end) return false;
  *v = x;
  return true;
}

// display-only statements and allocation hints (graphics, text, tiff output)
const char* const kIgnored[] = {"graphics", "graphic_iter", "graphic_delay", "color", "colour", "display_size",
                                "text_display", "frame_thickness", "frame_color", "grid_thickness", "grid_color",
                                "background_color", "tiff_name", "tiff_iter", "tiff_min", "tiff_max", "light",
                                "quit_at_end", "text_color", "line_style", "max_species", "max_names",
                                "max_surface", "max_cmd", "max_rxn", "max_mollist", "max_molecule_lists"};
// the same inside a start_surface block
const char* const kSurfIgnored[] = {"color", "colour", "thickness", "polygon", "shininess", "stipple",
                                    "max_panels", "stipple_factor"};

bool listed(const char* const* list, size_t n, const std::string& w) {
  for (size_t i = 0; i < n; i++)
    if (w == list[i]) return true;
  return false;
}

// strbegin(shortword, word, 0) (libSteve/string2.c:81-90): s is a non-empty
// leading part of word, compared without case
bool strbegin_ci(const std::string& s, const char* word) {
  size_t i = 0;
  for (; i < s.size(); i++)
    if (word[i] == '\0' || tolower((unsigned char)s[i]) != tolower((unsigned char)word[i])) return false;
  return i > 0;
}

// molstring2ms (smolmolec.c:75-89): state names
bool state_of(const std::string& s, MolecState* ms) {
  if (s == "soln" || s == "solution" || s == "fsoln" || s == "aq") *ms = MSsoln;
  else if (s == "front") *ms = MSfront;
  else if (s == "back") *ms = MSback;
  else if (s == "up") *ms = MSup;
  else if (s == "down") *ms = MSdown;
  else if (s == "bsoln") *ms = MSbsoln;
  else if (s == "all") *ms = MSall;
  else return false;
  return true;
}

// "name" or "name(state)" (molstring2pattern): the state defaults to solution
bool species_state(const std::string& tok, std::string* name, MolecState* ms) {
  *ms = MSsoln;
  if (!tok.empty() && tok.back() == ')') {
    size_t o = tok.rfind('(');
    if (o == std::string::npos) return false;
    if (!state_of(tok.substr(o + 1, tok.size() - o - 2), ms)) return false;
    *name = tok.substr(0, o);
  } else {
    *name = tok;
  }
  return !name->empty();
}

// surfstring2face (smolsurface.c:86-93): any leading part of the word, any case
bool face_of(const std::string& s, PanelFace* f) {
  if (strbegin_ci(s, "front")) *f = PFfront;
  else if (strbegin_ci(s, "back")) *f = PFback;
  else if (strbegin_ci(s, "all") || strbegin_ci(s, "both")) *f = PFboth;
  else return false;
  return true;
}

// surfstring2act (smolsurface.c:108-125), tested in the reference's order
bool action_of(const std::string& s, SrfAction* a) {
  if (strbegin_ci(s, "reflect")) *a = SAreflect;
  else if (strbegin_ci(s, "transmit")) *a = SAtrans;
  else if (strbegin_ci(s, "absorb")) *a = SAabsorb;
  else if (strbegin_ci(s, "jump")) *a = SAjump;
  else if (s == "periodic") *a = SAjump;
  else if (s == "port") *a = SAport;
  else if (strbegin_ci(s, "multiple")) *a = SAmult;
  else if (strbegin_ci(s, "no")) *a = SAno;
  else if (strbegin_ci(s, "adsorb")) *a = SAadsorb;
  else if (strbegin_ci(s, "revdes")) *a = SArevdes;
  else if (strbegin_ci(s, "irrevdes")) *a = SAirrevdes;
  else if (strbegin_ci(s, "flip")) *a = SAflip;
  else return false;
  return true;
}

// rxnstring2rp (smolreact.c:79-106): one-letter codes and full names; "b"
// is unbindrad, bounce is only spelled out
bool placement_of(const std::string& s, RevParam* rp) {
  static const std::map<std::string, RevParam> m = {
      {"i", RPirrev},         {"a", RPconfspread},    {"p", RPpgem},          {"x", RPpgemmax},
      {"r", RPratio},         {"b", RPunbindrad},     {"q", RPpgem2},         {"y", RPpgemmax2},
      {"s", RPratio2},        {"o", RPoffset},        {"f", RPfixed},         {"irrev", RPirrev},
      {"confspread", RPconfspread}, {"bounce", RPbounce}, {"pgem", RPpgem},   {"pgemmax", RPpgemmax},
      {"ratio", RPratio},     {"unbindrad", RPunbindrad}, {"pgem2", RPpgem2}, {"pgemmax2", RPpgemmax2},
      {"ratio2", RPratio2},   {"offset", RPoffset},   {"fixed", RPfixed}};
  auto it = m.find(s);
  if (it == m.end()) return false;
  *rp = it->second;
  return true;
}

// the last error of a libsmoldyn call, when it failed
bool failed(ErrorCode e, const char* what, std::string* err) {
  if (e >= ECwarning) return false;
  char fn[512] = {0}, msg[512] = {0};
  smolGetError(fn, msg, 0);
  *err = std::string(what) + ": " + msg;
  return true;
}

int surface_statement(simptr sim, const std::string& srf, const std::string& word, const std::vector<std::string>& w,
                      std::string* cursurf, std::string* err) {
  auto fail = [&](const std::string& m) {
    *err = m;
    return 1;
  };
  if (word == "end_surface") {
    cursurf->clear();
    return 0;
  }
  if (listed(kSurfIgnored, sizeof(kSurfIgnored) / sizeof(kSurfIgnored[0]), word)) return 0;
  if (word == "action") {
    // new format: species face action [new species]; old: face species action
    if (w.size() < 3) return fail("action format: species face action [new_species]");
    // the new format when its species is known and its face readable, else
    // the old one (smolsurface.c surfreadstring "action")
    auto known = [&](const std::string& t) {
      std::string nm;
      MolecState m;
      return species_state(t, &nm, &m) && (nm == "all" || smolGetSpeciesIndexNT(sim, nm.c_str()) > 0);
    };
    PanelFace f;
    std::string sp = w[0];
    if (!face_of(w[1], &f) || !known(w[0])) {
      PanelFace f0;
      const bool face0 = face_of(w[0], &f0), known1 = known(w[1]);
      if (face0 || known1) {
        if (!face0) return fail("in action, face name needs to be 'front', 'back', or 'both'");
        if (!known1) return fail("in action, molecule name not recognized");
        f = f0;
        sp = w[1];
      } else {
        if (!face_of(w[1], &f)) return fail("in action, face name needs to be 'front', 'back', or 'both'");
        return fail("in action, molecule name not recognized");
      }
    }
    std::string name;
    MolecState ms;
    if (!species_state(sp, &name, &ms)) return fail("action: species not readable");
    SrfAction a;
    if (!action_of(w[2], &a)) return fail("action '" + w[2] + "' is not run on the device path");
    const char* nsp = w.size() > 3 ? w[3].c_str() : nullptr;
    return failed(smolSetSurfaceAction(sim, srf.c_str(), f, name.c_str(), ms, a, nsp), "action", err) ? 1 : 0;
  }
  if (word == "rate" || word == "rate_internal") {
    if (w.size() < 4) return fail(word + " format: species(state) state1 state2 value [new_species]");
    std::string name;
    MolecState ms, ms1, ms2;
    double v;
    if (!species_state(w[0], &name, &ms)) return fail(word + ": species not readable");
    if (!state_of(w[1], &ms1) || !state_of(w[2], &ms2)) return fail(word + ": state not recognized");
    if (!num(w[3], &v)) return fail(word + ": value not readable");
    const char* nsp = w.size() > 4 ? w[4].c_str() : nullptr;
    return failed(smolSetSurfaceRate(sim, srf.c_str(), name.c_str(), ms, ms1, ms2, v, nsp, word == "rate_internal"),
                  word.c_str(), err)
               ? 1
               : 0;
  }
  if (word == "panel") {
    if (w.empty()) return fail("panel format: shape [axis] parameters [name]");
    const std::string sh = w[0];
    PanelShape ps;
    size_t npar;
    // surfstring2ps (smolsurface.c:143-155): leading parts, any case
    if (strbegin_ci(sh, "rectangle")) ps = PSrect, npar = 5;
    else if (strbegin_ci(sh, "triangle")) ps = PStri, npar = 9;
    else if (strbegin_ci(sh, "sphere")) ps = PSsph, npar = 6;
    else if (strbegin_ci(sh, "cylinder")) ps = PScyl, npar = 9;
    else if (strbegin_ci(sh, "hemisphere")) ps = PShemi, npar = 9;
    else if (strbegin_ci(sh, "disk")) ps = PSdisk, npar = 8;
    else return fail("panel shape '" + sh + "' not recognized");
    size_t k = 1;
    std::string axis;
    if (ps == PSrect) {
      if (w.size() < 2) return fail("panel rect needs an axis");
      axis = w[1];
      k = 2;
    }
    std::vector<double> prm;
    std::string pname;
    for (; k < w.size(); k++) {
      double v;
      if (num(w[k], &v)) prm.push_back(v);
      else pname = w[k];
    }
    if ((ps == PSsph || ps == PSdisk) && prm.size() + 2 == npar) prm.push_back(20);  // drawing slices/stacks default
    if (ps == PSsph && prm.size() + 1 == npar) prm.push_back(20);
    if (ps == PSdisk && prm.size() + 1 == npar) prm.push_back(20);
    if (prm.size() != npar) return fail("panel " + sh + ": wrong number of parameters");
    prm.resize(16, 0.0);
    return failed(smolAddPanel(sim, srf.c_str(), ps, pname.empty() ? nullptr : pname.c_str(),
                               axis.empty() ? nullptr : axis.c_str(), prm.data()),
                  "panel", err)
               ? 1
               : 0;
  }
  if (word == "jump") {
    if (w.size() < 5) return fail("jump format: panel1 face1 -> panel2 face2");
    PanelFace f1, f2;
    if (!face_of(w[1], &f1) || !face_of(w[4], &f2)) return fail("jump: face not recognized");
    const int bidir = w[2] == "<->" ? 1 : 0;
    if (!bidir && w[2] != "->") return fail("jump: arrow not recognized");
    return failed(smolSetPanelJump(sim, srf.c_str(), w[0].c_str(), f1, w[3].c_str(), f2, bidir), "jump", err) ? 1
                                                                                                                   : 0;
  }
  return fail("surface statement '" + word + "' is not run on the device path");
}

// reaction [name] r1 [+ r2] -> p1 [+ p2 ...] [rate]; <-> adds name+"rev"
// (rxnparsereaction, smolreact.c:2411-2600)
int reaction_statement(simptr sim, const std::vector<std::string>& w, std::string* err) {
  auto fail = [&](const std::string& m) {
    *err = m;
    return 1;
  };
  if (w.size() < 3) return fail("reaction format: name reactants -> products [rate]");
  if (w[0].find('=') != std::string::npos)
    return fail("reactions confined to a compartment or surface are not run on the device path");
  std::string rname = w[0];
  size_t k = 1;
  std::vector<std::pair<std::string, MolecState>> rct, prd;
  bool bidir = false;
  for (;;) {  // reactants
    if (k + 1 >= w.size()) return fail("reaction: missing arrow");
    if (w[k] != "0") {
      std::string n;
      MolecState ms;
      if (!species_state(w[k], &n, &ms)) return fail("reaction: reactant not readable");
      rct.push_back({n, ms});
    }
    const std::string& sym = w[k + 1];
    k += 2;
    if (sym == "->") break;
    if (sym == "<->") {
      bidir = true;
      break;
    }
    if (sym != "+") return fail("reaction: illegal symbol in reactant list");
  }
  for (;;) {  // products
    if (k >= w.size()) return fail("reaction: missing products");
    if (w[k] != "0") {
      std::string n;
      MolecState ms;
      if (!species_state(w[k], &n, &ms)) return fail("reaction: product not readable");
      prd.push_back({n, ms});
    }
    k++;
    if (k < w.size() && w[k] == "+") {
      k++;
      continue;
    }
    break;
  }
  if (rct.size() > 2) return fail("reaction: more than two reactants");
  double rates[2] = {-1, -1};
  for (int j = 0; j < 2 && k < w.size(); j++, k++)
    if (!num(w[k], &rates[j])) return fail("failed to read reaction rate");
  if (k < w.size()) return fail("unexpected text following reaction");
  auto add = [&](const std::string& nm, const std::vector<std::pair<std::string, MolecState>>& r,
                 const std::vector<std::pair<std::string, MolecState>>& p, double rate) {
    std::vector<const char*> pn;
    std::vector<MolecState> ps;
    for (auto& x : p) {
      pn.push_back(x.first.c_str());
      ps.push_back(x.second);
    }
    const char* r1 = r.size() > 0 ? r[0].first.c_str() : nullptr;
    const char* r2 = r.size() > 1 ? r[1].first.c_str() : nullptr;
    MolecState s1 = r.size() > 0 ? r[0].second : MSsoln, s2 = r.size() > 1 ? r[1].second : MSsoln;
    return smolAddReaction(sim, nm.c_str(), r1, s1, r2, s2, (int)p.size(), pn.data(), ps.data(), rate);
  };
  if (!bidir) return failed(add(rname, rct, prd, rates[0]), "reaction", err) ? 1 : 0;
  if (failed(add(rname + "fwd", rct, prd, rates[0]), "reaction", err)) return 1;
  return failed(add(rname + "rev", prd, rct, rates[1]), "reaction", err) ? 1 : 0;
}

}  // namespace

int config_statement(simptr sim, const std::string& word, const std::string& rest, std::string* cursurf,
                     std::string* err) {
  const std::vector<std::string> w = split(rest);
  auto fail = [&](const std::string& m) {
    *err = m;
    return 1;
  };
  if (!cursurf->empty()) return surface_statement(sim, *cursurf, word, w, cursurf, err);
  if (listed(kIgnored, sizeof(kIgnored) / sizeof(kIgnored[0]), word)) return 0;
  double v;
  long n;
  if (word == "random_seed" || word == "rand_seed") {
    if (w.size() != 1) return fail("random_seed needs to be 'time' or an integer");
    if (w[0] == "time") n = (long)time(nullptr);
    else if (!inum(w[0], &n)) return fail("random_seed needs to be 'time' or an integer");
    return failed(smolSetRandomSeed(sim, n), word.c_str(), err) ? 1 : 0;
  }
  if (word == "time_start" || word == "time_stop" || word == "time_step" || word == "time_now") {
    if (w.size() != 1 || !num(w[0], &v)) return fail(word + " needs to be a number");
    if (word == "time_start")  // simsettime codes 0 and 1: the clock and the start time
      return (failed(smolSetTimeStart(sim, v), word.c_str(), err) || failed(smolSetTimeNow(sim, v), word.c_str(), err))
                 ? 1
                 : 0;
    if (word == "time_stop") return failed(smolSetTimeStop(sim, v), word.c_str(), err) ? 1 : 0;
    if (word == "time_step") return failed(smolSetTimeStep(sim, v), word.c_str(), err) ? 1 : 0;
    return failed(smolSetTimeNow(sim, v), word.c_str(), err) ? 1 : 0;
  }
  if (word == "species" || word == "names" || word == "name") {
    for (auto& s : w)
      if (failed(smolAddSpecies(sim, s.c_str(), nullptr), word.c_str(), err)) return 1;
    return 0;
  }
  if (word == "difc") {
    if (w.size() != 2 || !num(w[1], &v)) return fail("difc format: species[(state)] value");
    std::string name;
    MolecState ms;
    if (!species_state(w[0], &name, &ms)) return fail("difc: species not readable");
    if (name == "all") {
      for (auto& s : sim_species_names(sim))
        if (failed(smolSetSpeciesMobility(sim, s.c_str(), ms, v, nullptr, nullptr), "difc", err)) return 1;
      return 0;
    }
    return failed(smolSetSpeciesMobility(sim, name.c_str(), ms, v, nullptr, nullptr), "difc", err) ? 1 : 0;
  }
  if (word == "molecule_lists") {
    for (auto& s : w)
      if (failed(smolAddMolList(sim, s.c_str()), word.c_str(), err)) return 1;
    return 0;
  }
  if (word == "mol_list") {
    if (w.size() != 2) return fail("mol_list format: species[(state)] list_name");
    std::string name;
    MolecState ms;
    if (!species_state(w[0], &name, &ms)) return fail("mol_list: species not readable");
    return failed(smolSetMolList(sim, name.c_str(), ms, w[1].c_str()), word.c_str(), err) ? 1 : 0;
  }
  if (word == "molperbox" || word == "boxsize") {
    if (w.size() != 1 || !num(w[0], &v)) return fail(word + " needs to be a number");
    return failed(smolSetPartitions(sim, word.c_str(), v), word.c_str(), err) ? 1 : 0;
  }
  if (word == "accuracy") {
    if (w.size() != 1 || !num(w[0], &v)) return fail("accuracy needs to be a number");
    sim_set_accuracy(sim, v);
    return 0;
  }
  if (word == "max_mol") {
    if (w.size() != 1 || !inum(w[0], &n)) return fail("max_mol needs to be an integer");
    return failed(smolSetMaxMolecules(sim, (int)n), word.c_str(), err) ? 1 : 0;
  }
  if (word == "mol") {
    // mol n species position...: u (uniform), lo-hi (range) or a value per axis
    if (w.size() < 2 || !inum(w[0], &n) || n < 0) return fail("mol format: number name position_vector");
    std::string name;
    MolecState ms;
    if (!species_state(w[1], &name, &ms) || ms != MSsoln) return fail("mol: species not readable or not in solution");
    double lo[3], hi[3];
    sim_bounds(sim, lo, hi);  // 'u': the system corners (systemcorners)
    const int dim = sim_dim(sim);
    if ((int)w.size() < 2 + dim) return fail("insufficient position data for mol");
    if ((int)w.size() > 2 + dim) return fail("unexpected text following mol");
    for (int d = 0; d < dim; d++) {
      const std::string& t = w[2 + (size_t)d];
      if (t[0] == 'u') continue;
      double a, b;
      if (sscanf(t.c_str(), "%lg-%lg", &a, &b) == 2) {
        lo[d] = a;
        hi[d] = b;
      } else if (num(t, &a)) {
        lo[d] = hi[d] = a;
      } else {
        return fail("cannot read position value for mol");
      }
    }
    double* plo = lo;
    double* phi = hi;
    return failed(smolAddSolutionMolecules(sim, name.c_str(), (int)n, plo, phi), "mol", err) ? 1 : 0;
  }
  if (word == "reaction") return reaction_statement(sim, w, err);
  if (word == "reaction_rate" || word == "binding_radius" || word == "reaction_probability") {
    if (w.size() != 2 || !num(w[1], &v)) return fail(word + " format: reaction value");
    const int type = word == "reaction_rate" ? 0 : word == "binding_radius" ? 1 : 2;
    return failed(smolSetReactionRate(sim, w[0].c_str(), v, type), word.c_str(), err) ? 1 : 0;
  }
  if (word == "confspread_radius") {
    if (w.size() != 2 || !num(w[1], &v)) return fail("confspread_radius format: rname radius");
    if (failed(smolSetReactionProducts(sim, w[0].c_str(), RPconfspread, 0, nullptr, nullptr), word.c_str(), err))
      return 1;
    return failed(smolSetReactionRate(sim, w[0].c_str(), v, 1), word.c_str(), err) ? 1 : 0;
  }
  if (word == "reaction_log") {
    // reaction_log filename rxnname|all serial_numbers|all (smolsim.cpp:2016-2040)
    if (w.size() < 3) return fail("reaction_log format: filename rxnname serial_numbers");
    std::vector<long> sn;
    int n = -1;
    if (w[2] != "all") {
      for (size_t k = 2; k < w.size(); k++) {
        long v;
        if (!inum(w[k], &v)) return fail("unable to read serial numbers");
        sn.push_back(v);
      }
      n = (int)sn.size();
    } else if (w.size() > 3) {
      return fail("unexpected text following reaction_log");
    }
    return failed(smolamdSetReactionLog(sim, w[0].c_str(), w[1].c_str(), sn.data(), n, 1), word.c_str(), err) ? 1
                                                                                                               : 0;
  }
  if (word == "product_placement") {
    // product_placement rname type [parameter] / offset|fixed product x y z
    if (w.size() < 2) return fail("product_placement format: rname type parameters");
    RevParam rp;
    if (!placement_of(w[1], &rp)) return fail("unrecognized product placement '" + w[1] + "'");
    int nprod = 0;
    const int order = sim_reaction_shape(sim, w[0].c_str(), &nprod);
    if (order < 0) return fail("unrecognized reaction name");
    if (rp == RPbounce && (order != 2 || nprod < 2))
      return fail("bounce can only be used with two reactants and at least two products");
    double prm = 0;
    if (rp == RPoffset || rp == RPfixed) {
      // the product by species name or as product_N (1-based), then its offset vector
      const int dim = sim_dim(sim);
      if ((int)w.size() != 3 + dim) return fail("product_placement format: rname type product position_vector");
      double pos[3] = {0, 0, 0};
      for (int d = 0; d < dim; d++)
        if (!num(w[3 + (size_t)d], &pos[d])) return fail("insufficient data for position vector for product_placement");
      std::string name;
      MolecState ms;
      if (w[2].compare(0, 8, "product_") == 0) {
        long k;
        if (!inum(w[2].substr(8), &k)) return fail("error reading product number");
        name = sim_product_name(sim, w[0].c_str(), (int)k - 1);
        if (name.empty()) return fail("illegal product number");
      } else if (!species_state(w[2], &name, &ms)) {
        return fail("product_placement: product not readable");
      }
      return failed(smolSetReactionProducts(sim, w[0].c_str(), rp, 0, name.c_str(), pos), word.c_str(), err) ? 1
                                                                                                              : 0;
    }
    if (rp == RPbounce && w.size() == 2) prm = -1;  // default bounce method
    else if (rp != RPirrev && rp != RPnone && rp != RPconfspread && (w.size() < 3 || !num(w[2], &prm)))
      return fail("error reading parameter in product_placement");
    return failed(smolSetReactionProducts(sim, w[0].c_str(), rp, prm, nullptr, nullptr), word.c_str(), err) ? 1 : 0;
  }
  if (word == "start_surface") {
    if (w.size() != 1) return fail("start_surface format: name");
    if (smolGetSurfaceIndexNT(sim, w[0].c_str()) < 0 &&
        failed(smolAddSurface(sim, w[0].c_str()), word.c_str(), err))
      return 1;
    *cursurf = w[0];
    return 0;
  }
  if (word == "output_root") {
    if (w.size() != 1) return fail("output_root format: path");
    return failed(smolSetOutputPath(sim, w[0].c_str()), word.c_str(), err) ? 1 : 0;
  }
  if (word == "output_files" || word == "output_file" || word == "append_files") {
    for (auto& s : w) {
      std::string f = s;
      if (failed(smolAddOutputFile(sim, &f[0], -1, word == "append_files"), word.c_str(), err)) return 1;
    }
    return 0;
  }
  if (word == "output_data") {
    for (auto& s : w) {
      std::string d = s;
      if (failed(smolAddOutputData(sim, &d[0]), word.c_str(), err)) return 1;
    }
    return 0;
  }
  if (word == "output_precision") {
    if (w.size() != 1 || !inum(w[0], &n)) return fail("format for output_precision: value");
    sim_set_output_precision(sim, (int)n);
    return 0;
  }
  if (word == "output_format") {
    if (w.size() != 1 || sim_set_output_format(sim, w[0].c_str())) return fail("output_format value not recognized");
    return 0;
  }
  if (word == "output_file_number") {
    if (w.size() != 2 || !inum(w[1], &n) || n < 0) return fail("format for output_file_number: filename number");
    if (sim_set_file_suffix(sim, w[0].c_str(), (int)n)) return fail("error setting output_file_number");
    return 0;
  }
  if (word == "cmd") {
    std::string line = rest;
    return failed(smolAddCommandFromString(sim, &line[0]), "cmd", err) ? 1 : 0;
  }
  return fail("statement '" + word + "' is not run on the device path");
}

namespace {

struct Line {
  std::string file;
  int lineno;
  std::string word, rest;
};

// Read a file into statements: comments (# to the end of the line, /* */
// blocks), define substitution, ifdefine / ifundefine / else / endif,
// read_file, end_file (smolsim.cpp loadsim and simpreloadfile).
bool read_lines(const std::string& dir, const std::string& name, std::map<std::string, std::string>& defs,
                std::vector<Line>* out, std::string* err, int depth) {
  if (depth > 16) {
    *err = "read_file nested too deep";
    return false;
  }
  std::ifstream f(dir + name);
  if (!f) {
    *err = "cannot open file " + dir + name;
    return false;
  }
  std::string raw;
  int lineno = 0;
  bool incomment = false;
  std::vector<bool> skip{false};
  while (std::getline(f, raw)) {
    lineno++;
    std::string s;
    for (size_t i = 0; i < raw.size(); i++) {  // strip comments
      if (incomment) {
        if (raw.compare(i, 2, "*/") == 0) {
          incomment = false;
          i++;
        }
        continue;
      }
      if (raw.compare(i, 2, "/*") == 0) {
        incomment = true;
        i++;
        continue;
      }
      if (raw[i] == '#') break;
      s += raw[i];
    }
    std::vector<std::string> w = split(s);
    if (w.empty()) continue;
    const std::string& k = w[0];
    if (k == "ifdefine") {
      skip.push_back(skip.back() || w.size() < 2 || !defs.count(w[1]));
      continue;
    }
    if (k == "ifundefine") {
      skip.push_back(skip.back() || (w.size() > 1 && defs.count(w[1])));
      continue;
    }
    if (k == "else") {
      if (skip.size() < 2) {
        *err = name + ":" + std::to_string(lineno) + ": else without ifdefine";
        return false;
      }
      bool top = skip.back();
      skip.pop_back();
      skip.push_back(skip.back() ? true : !top);
      continue;
    }
    if (k == "endif") {
      if (skip.size() < 2) {
        *err = name + ":" + std::to_string(lineno) + ": endif without ifdefine";
        return false;
      }
      skip.pop_back();
      continue;
    }
    if (skip.back()) continue;
    if (k == "define" || k == "define_global") {
      std::string v;
      for (size_t i = 2; i < w.size(); i++) v += (i > 2 ? " " : "") + w[i];
      if (w.size() > 1) defs[w[1]] = v;
      continue;
    }
    if (k == "undefine") {
      if (w.size() > 1) defs.erase(w[1]);
      continue;
    }
    if (k == "end_file") break;
    for (auto& t : w) {  // defined names are replaced word by word
      auto it = defs.find(t);
      if (it != defs.end()) t = it->second;
    }
    if (w[0] == "read_file") {
      if (w.size() < 2 || !read_lines(dir, w[1], defs, out, err, depth + 1)) {
        if (err->empty()) *err = name + ":" + std::to_string(lineno) + ": read_file needs a file name";
        return false;
      }
      continue;
    }
    Line L;
    L.file = name;
    L.lineno = lineno;
    L.word = w[0];
    for (size_t i = 1; i < w.size(); i++) L.rest += (i > 1 ? " " : "") + w[i];
    out->push_back(L);
  }
  if (skip.size() != 1) {
    *err = name + ": ifdefine without endif";
    return false;
  }
  return true;
}

int axis_of(const std::string& s) {
  if (s == "0" || s == "x") return 0;
  if (s == "1" || s == "y") return 1;
  if (s == "2" || s == "z") return 2;
  return -1;
}

}  // namespace

// loadsim: a new simulation from dim and the boundary statements, then every
// other statement in file order.  *sim is kept when given (its dimension and
// bounds must match the file's).
int config_load(const char* filepath, const char* filename, simptr* simp, std::string* err) {
  std::string dir = filepath ? filepath : "";
  if (!dir.empty() && dir.back() != '/') dir += '/';
  std::map<std::string, std::string> defs;
  std::vector<Line> lines;
  if (!read_lines(dir, filename, defs, &lines, err, 0)) return 1;
  int dim = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  char wt[3][2] = {{'t', 't'}, {'t', 't'}, {'t', 't'}};
  bool have[3][2] = {{false, false}, {false, false}, {false, false}};
  auto where = [&](const Line& L) { return L.file + ":" + std::to_string(L.lineno) + ": "; };
  for (auto& L : lines) {
    std::vector<std::string> w = split(L.rest);
    if (L.word == "dim") {
      long d;
      if (w.size() != 1 || !inum(w[0], &d) || d < 1 || d > 3) {
        *err = where(L) + "dim needs to be 1, 2 or 3";
        return 1;
      }
      dim = (int)d;
    } else if (L.word == "boundaries") {
      int d = w.empty() ? -1 : axis_of(w[0]);
      double a, b;
      if (d < 0 || w.size() < 3 || !num(w[1], &a) || !num(w[2], &b) || !(a < b)) {
        *err = where(L) + "boundaries format: dimension position position [type]";
        return 1;
      }
      lo[d] = a;
      hi[d] = b;
      const char t = w.size() > 3 ? w[3][0] : 't';
      wt[d][0] = wt[d][1] = t;
      have[d][0] = have[d][1] = true;
    } else if (L.word == "low_wall" || L.word == "high_wall") {
      int d = w.empty() ? -1 : axis_of(w[0]);
      double a;
      if (d < 0 || w.size() < 3 || !num(w[1], &a)) {
        *err = where(L) + L.word + " format: dimension position type";
        return 1;
      }
      const int side = L.word == "high_wall";
      (side ? hi : lo)[d] = a;
      wt[d][side] = w[2][0];
      have[d][side] = true;
    }
  }
  if (dim == 0) {
    *err = "the file does not set dim";
    return 1;
  }
  for (int d = 0; d < dim; d++)
    if (!have[d][0] || !have[d][1]) {
      *err = "boundaries are missing for dimension " + std::to_string(d);
      return 1;
    }
  simptr sim = *simp;
  const bool created = !sim;
  if (created) {
    sim = smolNewSim(dim, lo, hi);
    if (!sim) {
      char fn[512], msg[512];
      smolGetError(fn, msg, 0);
      *err = std::string("smolNewSim: ") + msg;
      return 1;
    }
  }
  for (int d = 0; d < dim; d++)
    for (int side = 0; side < 2; side++)
      if (failed(smolSetBoundaryType(sim, d, side, wt[d][side]), "boundaries", err)) {
        if (created) smolFreeSim(sim);
        return 1;
      }
  std::string cursurf;
  for (auto& L : lines) {
    if (L.word == "dim" || L.word == "boundaries" || L.word == "low_wall" || L.word == "high_wall") continue;
    if (config_statement(sim, L.word, L.rest, &cursurf, err)) {
      *err = where(L) + *err;
      if (created) smolFreeSim(sim);
      return 1;
    }
  }
  *simp = sim;
  return 0;
}

}  // namespace smolamd
