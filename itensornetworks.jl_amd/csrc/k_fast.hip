// Specialised sweep pipelines (placeholder: none yet).
#include "itnbp_ops.hpp"

bool fast_path_available(itnbp_net*, int, int) { return false; }
const char* fast_path_name(itnbp_net*, int) { return "none"; }
int fast_parallel_sweep(itnbp_net*, const std::vector<int>&, int, const void*, int, void*, void*,
                        MsgStat*, int, bool* done) {
  *done = false;
  return ITNBP_OK;
}
void release_fast_state(itnbp_net*) {}
void invalidate_fast_state(itnbp_net*) {}
