// FP32 instantiations of the fused hex operator kernel (FP32 storage and arithmetic).
#include "hex_plan.hpp"

namespace mi355 {
ApplyFn pick_apply_f32(int P, int Q, int num_comp, int qf) { return pick_apply<float>(P, Q, num_comp, qf); }
}  // namespace mi355
