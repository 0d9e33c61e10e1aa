// ref_world.cpp — TEST INFRASTRUCTURE ONLY. A C entry point for the reference's own world-cell
// arithmetic (core/math/include/core/math/world.h, the float32 `relative_axis` its shaders restate
// as scene.slang's instance_from_eye), compiled from the reference's header where it lies
// (oracle/Makefile, target `ref`) into oracle/_ref/. The tests hold the oracle's restatement to it.
#include <core/math/world.h>

#include <cstdint>

extern "C" float ge_ref_relative_axis(int32_t cell, float local, int32_t eye_cell, float eye_local,
                                      float eye_residual) {
  return engine::relative_axis(cell, local, eye_cell, eye_local, eye_residual);
}
