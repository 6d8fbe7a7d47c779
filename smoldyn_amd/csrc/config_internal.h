// Simulation settings that configuration statements reach but the libsmoldyn
// C API does not (capi.cpp), and the statement reader (config.cpp).
#pragma once

#include <string>
#include <vector>

#include "../../include/libsmoldyn_amd.h"

namespace smolamd {

// dimension, system corners and species names (index 1 on)
int sim_dim(simptr sim);
void sim_bounds(simptr sim, double* lo, double* hi);
std::vector<std::string> sim_species_names(simptr sim);
// species of product prd of a reaction ("" when there is none)
std::string sim_product_name(simptr sim, const char* reaction, int prd);
// order of a reaction (-1 when there is none) and its product count
int sim_reaction_shape(simptr sim, const char* reaction, int* nprod);
// accuracy (smolsim.cpp "accuracy" statement)
void sim_set_accuracy(simptr sim, double accuracy);
// output_precision, output_format (ssv / csv), output_file_number
void sim_set_output_precision(simptr sim, int precision);
int sim_set_output_format(simptr sim, const char* format);
int sim_set_file_suffix(simptr sim, const char* name, int suffix);

// Read a configuration file (filepath + filename) into *sim, creating it from
// the file's dim and boundaries when *sim is NULL.  Returns 0, or nonzero with
// "file:line: message" in err.
int config_load(const char* filepath, const char* filename, simptr* sim, std::string* err);

// Apply one configuration statement (word + the rest of the line) to sim.
// cursurf names the surface of an open start_surface block ("" outside).
// Returns 0, or nonzero with a message in err.
int config_statement(simptr sim, const std::string& word, const std::string& rest, std::string* cursurf,
                     std::string* err);

}  // namespace smolamd
