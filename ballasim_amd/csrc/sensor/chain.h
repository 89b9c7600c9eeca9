// The launch chain of one advance call over one camera or a rig of cameras: zero the counters, phase
// A over every segment of every camera, the heavy-first order of the joint pixel space, then one walk
// launch over all pixels (persistent, fed by a work queue, when the tuning's chunk is positive).
//
// The kernels read the cameras' descriptors and the call header from device memory, and their grids
// are sized by capacities, not by the call: so the chain is captured once into a hipGraph and every
// later call whose joint run and segment counts fit the captured capacities is one graph launch
// after two small host-to-device copies (the descriptors and the header, from pinned memory). A
// call that outgrows them, a new tuning or a new set of cameras re-captures.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "kconst.h"
#include "kernels.h"
#include "model.h"

namespace bsx {

struct ChainStats {
  uint64_t launches = 0;        // calls launched (graph replays and kernel-by-kernel launches)
  uint64_t graph_launches = 0;  // calls launched as one graph
  uint64_t captures = 0;        // graphs captured
  uint64_t walk_reruns = 0;     // walks re-run after an event-buffer overflow
};

class LaunchChain {
 public:
  LaunchChain();
  ~LaunchChain();
  LaunchChain(const LaunchChain&) = delete;
  LaunchChain& operator=(const LaunchChain&) = delete;

  void set_tuning(const WalkTuning& tu);
  const WalkTuning& tuning() const { return tu_; }
  void set_graph(bool on);
  bool graph() const { return graph_on_; }
  const ChainStats& stats() const { return stats_; }

  // Lays the cameras out in the joint pixel, run and segment spaces (sets base, run_base and
  // seg_base), uploads the descriptors and the call header and launches the whole chain on stream.
  void launch(std::vector<CameraDev>& cams, bool instant, int64_t t_end, hipStream_t stream);
  // After an overflow: the descriptors again (new event buffers), the counters and queue zeroed, and
  // the walk alone (phase A and the order of the last launch are reused).
  void rerun_walk(const std::vector<CameraDev>& cams, hipStream_t stream);
  // The walk alone with another tuning, after a launch of the same cameras (the autotuner's trials).
  void walk_with(const WalkTuning& tu, hipStream_t stream);

 private:
  ChainArgs args(const WalkTuning& tu) const;
  void upload(const std::vector<CameraDev>& cams, int64_t t_end, hipStream_t stream);
  void capture();
  void drop_graph();

  WalkTuning tu_{256, 8, 8, 0};
  bool graph_on_ = true;
  ChainStats stats_{};

  DeviceBuffer d_cams_, d_call_, d_key_, d_perm_, d_hist_, d_queue_;
  CameraDev* h_cams_ = nullptr;  // pinned staging of the descriptors and the header
  CallDev* h_call_ = nullptr;
  int h_cap_ = 0;
  hipEvent_t staged_ = nullptr;  // recorded after the last upload from the staging

  // The shape of the last launch (what a captured graph depends on).
  int n_cams_ = 0;
  uint32_t n_total_ = 0;
  int groups_ = 1;
  bool instant_ = true;
  bool noise_ = false;
  int64_t seg_capacity_ = 0, run_capacity_ = 0;
  int64_t t_end_ = 0;

  hipStream_t capture_stream_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool graph_valid_ = false;
};

}  // namespace bsx
