// sph_derive.hpp - host-side derivations of sph_params shared by the scene
// helpers and the device contexts.
#pragma once

#include "../../include/pibiti_sph.h"

namespace sph {

// Fill the derived boundary fields (wave_zmin_shift, pump_xs, pump_zs) and
// the physics bits from the raw fields.
void derive_shape(sph_params& p);

// SPH_ERR_INVALID_PARAMS for an unknown boundary type, effect or accelerator.
int check_shape(const sph_params& p);

// The per-step advance of the rotor angle (the wave phase) and of the second
// propeller's angle (no-op when rot_vel == 0).
void advance_motion(sph_params& p);

// Everything the reference's app changes before each step: the dye-clear
// count, then advance_motion.
void advance_frame(sph_params& p);

// True when advance_frame changes the parameters (a turning rotor or
// moving wave, or dye clears still pending).
bool changes_per_step(const sph_params& p);

}  // namespace sph
