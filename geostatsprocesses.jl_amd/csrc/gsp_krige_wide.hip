// gsp_krige_wide.hip — kriging systems of 49..64 neighbours in FP64
// (k_krige<64, FP64>), and the dispatch of the buckets above 32 neighbours to
// the translation units that hold their kernels (compiled apart because the
// fully unrolled column loops of these kernels dominate the build).
#include "gsp_krige.h"

namespace gsp {

void launch_krige_48(const PlanArgs& A, int precision, const int32_t* list, int64_t count,
                     hipStream_t s);
void launch_krige_64r(const PlanArgs& A, const int32_t* list, int64_t count, hipStream_t s);

void launch_krige_wide(int KB, const PlanArgs& A, int precision, const int32_t* list,
                       int64_t count, hipStream_t s) {
  if (KB == 48)
    launch_krige_48(A, precision, list, count, s);
  else if (precision == GSP_FP32_REFINE)
    launch_krige_64r(A, list, count, s);
  else
    launch_krige_prec<64, GSP_FP64>(A, list, count, s);
}

}  // namespace gsp
