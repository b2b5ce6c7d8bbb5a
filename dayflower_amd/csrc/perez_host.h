// perez_host.h — host-side construction of the Perez sun/sky record.
#pragma once
#include "device_types.h"

namespace dfr {

// Fills `p` for PerezLight(turbidity, sunDirectionWorldSpace)
// (M/scene/light/PerezLight.java set(...)).
void perez_build(DevPerez& p, float turbidity, const float sun_direction_world[3]);

// Fills `L` for ImageLight(image, AngleF.degrees(angle), scale)
// (M/scene/light/ImageLight.java:134-150, 544-576) over the image whose
// width x height RGB texels start at texels + texel_offset.
void image_light_build(DevImageLight& L, int width, int height, int64_t texel_offset, const float* texels,
                       float angle_degrees, float scale_u, float scale_v, float radius);
// ImageLight.doRadianceSky of an object-space direction.
void image_light_radiance(const DevImageLight& L, const float* texels, const float d[3], float rgb[3]);

}  // namespace dfr
