// Fused particle-kernel instantiations, part 2 of the profile-combination table
// (split over translation units so they compile in parallel).
#include "fused_entry.h"

namespace glc {

const ModelEntry* models_2(int* n) {
    static const ModelEntry table[] = {
        GL_ENTRY(SIE, SHR, 0, SE, 0, SE, 0),
        GL_ENTRY(EPL, SHR, TNFW, SE, 0, SE, 0),  // main deflector + one truncated-NFW subhalo
        GL_ENTRY(EPL, SHR, 0, CSE, 0, SE, 0),    // core-Sersic lens light
    };
    *n = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace glc
