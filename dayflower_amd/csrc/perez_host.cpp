// perez_host.cpp — PerezLight setup on the host.
#include "perez_host.h"

#include <cstring>

namespace dfr {

void perez_build(DevPerez& p, float turbidity, const float sun_direction_world[3]) {
    std::memset(&p, 0, sizeof(p));
    (void)turbidity;
    for (int k = 0; k < 3; k++) p.sun_dir_world[k] = sun_direction_world[k];
}

}  // namespace dfr
