// rd_camera.h — host setup of the perspective camera (product side).
#pragma once
#include "rd_types.h"
#include "rd_light.h"
#include "../../include/radiantium_hip.h"

namespace rd {

// PerspectiveCamera ctor + Update (Radiantium.Offline/Camera.cs:101-180):
// LookAtLeftHand (MathExt.cs:149-161) or the given camera-to-world matrix,
// clip = perspective * translate * scale, clipToCamera = inverse(clip).
// Returns false when a matrix is singular (the reference throws).  lt
// (optional) receives what SampleWi / Importance read (Camera.cs:172-179).
bool setup_camera(const rd_camera_desc& c, DevCamera* out, LtCamera* lt = nullptr);

// Matrix4x4.Invert by the adjugate of 2x2 minors; false when |det| < float.Epsilon.
bool invert4(const float* m, float* out);
void mul4(const float* a, const float* b, float* out);

}  // namespace rd
