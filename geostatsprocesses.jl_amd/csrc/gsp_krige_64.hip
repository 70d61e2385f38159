// gsp_krige_64.hip — kriging systems of 49..64 neighbours with FP32 factors
// and refinement (k_krige<64, FP32 + refinement>); a translation unit of its
// own so the build compiles this large unrolled kernel beside the rest.
#include "gsp_krige.h"

namespace gsp {

void launch_krige_64r(const PlanArgs& A, const int32_t* list, int64_t count, hipStream_t s) {
  launch_krige_prec<64, GSP_FP32_REFINE>(A, list, count, s);
}

}  // namespace gsp
