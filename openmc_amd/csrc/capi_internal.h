// What the model-query part of the C API (capi_model.cpp) needs from the
// driver's state (capi.cpp).
#pragma once

#include <string>

#include "host_model.h"

namespace omc {

// The loaded model, or null before openmc_init / after openmc_finalize.
HostModel* api_model();
// Changes whenever a model is loaded or released (tables derived from the
// model are rebuilt when it changes).
long long api_generation();
// Whether openmc_simulation_init has run and openmc_simulation_finalize not yet.
bool api_simulation_running();
// Zero one tally's SUM/SUM_SQ (device and host mirror) and restart its
// realization count (Tally::reset, src/tallies/tally.cpp).
void api_tally_reset(int index);
// Remembers the path of the last statepoint written (path_statepoint_c).
void api_note_statepoint(const std::string& path);
// Sets openmc_err_msg.
void api_set_errmsg(const std::string& msg);

} // namespace omc
