// pending.cpp -- entry points whose device paths are added in later files of this library.
#include "plan.hpp"

extern "C" {
int ettk_ccc_grad_batch(ettk_plan*, const float*, int, const ettk_particle*, int, const int*, const double*,
                        const double*, double*, double*) {
    ettk::set_error("ettk_ccc_grad_batch: not available in this build");
    return ETTK_ERR_UNSUPPORTED;
}
int ettk_avg_reset(ettk_plan*) { return ETTK_ERR_UNSUPPORTED; }
int ettk_avg_accumulate(ettk_plan*, const float*, int, const ettk_particle*, const float*, const double*) {
    return ETTK_ERR_UNSUPPORTED;
}
int ettk_avg_finalize(ettk_plan*, float*) { return ETTK_ERR_UNSUPPORTED; }
}
