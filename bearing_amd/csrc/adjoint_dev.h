// The adjoint form's device side as the C ABI's context calls it (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/bearing_mi355x.h"

namespace bm {
struct HeldDevice;
void held_free(HeldDevice* h);
int held_load(int device, const bm_held* held, HeldDevice** out, std::string& err);
int held_forward(HeldDevice* h, hipStream_t stream, const double* state0, const double* mass, const double* friction,
                 const double* restitution, double* state_out, std::string& err);
int held_backward(HeldDevice* h, hipStream_t stream, const double* state_bar, double* state0_bar, double* mass_bar,
                  double* friction_bar, double* restitution_bar, std::string& err);
void held_times(const HeldDevice* h, double* forward_ms, double* backward_ms);
int held_branches(HeldDevice* h, hipStream_t stream, uint64_t* per_world, std::string& err);
}  // namespace bm
