// so3.hip -- S^2 projection and SO(3) rotational search (filled in below the CCC path).
#include "plan.hpp"

namespace ettk {
struct So3State {};
}  // namespace ettk

namespace ettk {
void so3_state_destroy(ettk_plan* P) {
    delete P->so3;
    P->so3 = nullptr;
}
void so3_reference_set(ettk_plan* P) { (void)P; }
}  // namespace ettk

using namespace ettk;

extern "C" {
int ettk_project_batch(ettk_plan*, const float*, int, const ettk_particle*, float*) { return ETTK_ERR_UNSUPPORTED; }
int ettk_so3_correlate_batch(ettk_plan*, int, const double*, const double*, double*) { return ETTK_ERR_UNSUPPORTED; }
int ettk_align_batch(ettk_plan*, const float*, int, const ettk_particle*, int, ettk_result*) { return ETTK_ERR_UNSUPPORTED; }
int ettk_ccc_grad_batch(ettk_plan*, const float*, int, const ettk_particle*, int, const int*, const double*,
                        const double*, double*, double*) {
    return ETTK_ERR_UNSUPPORTED;
}
int ettk_avg_reset(ettk_plan*) { return ETTK_ERR_UNSUPPORTED; }
int ettk_avg_accumulate(ettk_plan*, const float*, int, const ettk_particle*, const float*, const double*) {
    return ETTK_ERR_UNSUPPORTED;
}
int ettk_avg_finalize(ettk_plan*, float*) { return ETTK_ERR_UNSUPPORTED; }
}
