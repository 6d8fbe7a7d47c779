// Process-wide rate surrogate consulted by the setup (surrogate_host.cpp).
#pragma once

namespace smolamd {
namespace surrogate {
bool enabled();
// Binding radius from the surrogate when the reference's rate function
// confirms it, else from the numerical inversion; *used tells which.
double binding_radius(double rate, double dt, double difc, double ratio, bool* used);
}  // namespace surrogate
}  // namespace smolamd
