// ORACLE — test infrastructure only (see rt_core.hpp).
// The scene handle behind the oracle's C ABI: a loaded scene and its prepared lights and media.
#pragma once
#include "rt_render.hpp"

namespace orc {
struct Handle {
  Scene scene;
  Prepared prep;
};
}  // namespace orc
