// 3D Laguerre-cell kernel (placeholder until the polyhedron kernel lands).
#include "internal.h"

namespace sdot_amd {
void run_cells3(sdot_pd *pd, CellArgs &a) {
    (void)pd;
    (void)a;
    throw StatusError{SDOT_NOT_SUPPORTED, "3D cells are not built yet"};
}
}  // namespace sdot_amd
