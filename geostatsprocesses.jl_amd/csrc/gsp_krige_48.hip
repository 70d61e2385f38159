// gsp_krige_48.hip — kriging systems of 33..48 neighbours (k_krige<48>), one
// wavefront per system; a translation unit of its own so the build compiles
// this large unrolled kernel beside the rest.
#include "gsp_krige.h"

namespace gsp {

void launch_krige_48(const PlanArgs& A, int precision, const int32_t* list, int64_t count,
                     hipStream_t s) {
  launch_krige<48>(A, precision, list, count, s);
}

}  // namespace gsp
