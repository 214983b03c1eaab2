// FP64 instantiations of the fused hex operator kernel.
#include "hex_plan.hpp"

namespace mi355 {
ApplyFn pick_apply_f64(int P, int Q, int num_comp, int qf) { return pick_apply<double>(P, Q, num_comp, qf); }
}  // namespace mi355
