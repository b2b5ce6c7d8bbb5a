// perez_host.h — host-side construction of the Perez sun/sky record.
#pragma once
#include "device_types.h"

namespace dfr {

// Fills `p` for PerezLight(turbidity, sunDirectionWorldSpace)
// (M/scene/light/PerezLight.java set(...)).
void perez_build(DevPerez& p, float turbidity, const float sun_direction_world[3]);

}  // namespace dfr
