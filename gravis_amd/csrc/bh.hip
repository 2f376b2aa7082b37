// bh.hip -- Barnes-Hut on gfx950 (placeholder until the linear octree lands).
#include "internal.h"

namespace gravis {

gravis_status bh_force(Ctx* c, const gravis_solver*, size_t, const double*, const double*,
                       const double*, const double*, const FinishArgs&) {
    return fail(c, GRAVIS_UNSUPPORTED, "Barnes-Hut is not built yet");
}

}  // namespace gravis
