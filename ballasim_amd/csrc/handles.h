// The opaque handles of the C ABI (include/bsx/*.h): each owns one object of the library.
#pragma once

namespace bsx {
class SensorModel;
class Rig;
class Proposer;
class Scene;
class LearnedSampler;
class OutputStage;
class EventStream;
}  // namespace bsx

struct bsx_sensor {
  bsx::SensorModel* model;
};
struct bsx_rig {
  bsx::Rig* rig;
};
struct bsx_mlp {
  bsx::Proposer* p;
};
struct bsx_scene {
  bsx::Scene* s;
};
struct bsx_learned_sampler {
  bsx::LearnedSampler* ls;
};
struct bsx_output {
  bsx::OutputStage* o;
};
struct bsx_event_stream {
  bsx::EventStream* s;
};
