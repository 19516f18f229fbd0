// Fused particle-kernel instantiations, part 1 of the profile-combination table
// (split over translation units so they compile in parallel).
#include "fused_entry.h"

namespace glc {

const ModelEntry* models_1(int* n) {
    static const ModelEntry table[] = {
        GL_ENTRY(SIS, 0, 0, 0, 0, SE, 0),
        GL_ENTRY(SIS, SHR, 0, SE, 0, SE, 0),
        GL_ENTRY(EPL, SHR, 0, SE, 0, SC, 0),
        GL_ENTRY(EPL, SHR, 0, SE, SE, SE, 0),
    };
    *n = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace glc
