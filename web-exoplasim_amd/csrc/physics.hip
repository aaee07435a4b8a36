// Column physics of the ExoPlaSim step (gridpointd's parameterizations).
#include "engine.h"

namespace exo {

void enqueue_physics(Engine&, const std::vector<int>&) {}

}  // namespace exo
