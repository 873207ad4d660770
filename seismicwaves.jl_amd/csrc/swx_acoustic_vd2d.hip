// 2D acoustic variable-density engine (staggered pressure-velocity).
#include "swx_internal.hpp"

namespace swx {
std::unique_ptr<Engine> make_acoustic_vd2d(PlanCore &core) {
  (void)core;
  throw Error(SWX_EUNSUPPORTED, "acoustic variable-density engine not built yet");
}
}  // namespace swx
