// The launch chain of an advance call (one camera or a rig): layout of the joint spaces, descriptor
// upload, graph capture and replay.
#include "chain.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace bsx {

namespace {

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace

LaunchChain::LaunchChain() { hip_check(hipEventCreateWithFlags(&staged_, hipEventDisableTiming), "event create"); }

LaunchChain::~LaunchChain() {
  drop_graph();
  if (capture_stream_ != nullptr) (void)hipStreamDestroy(capture_stream_);
  if (h_cams_ != nullptr) (void)hipHostFree(h_cams_);
  if (h_call_ != nullptr) (void)hipHostFree(h_call_);
  if (staged_ != nullptr) (void)hipEventDestroy(staged_);
}

void LaunchChain::drop_graph() {
  if (exec_ != nullptr) (void)hipGraphExecDestroy(exec_);
  if (graph_ != nullptr) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  graph_valid_ = false;
}

void LaunchChain::set_tuning(const WalkTuning& tu) {
  if (tu.block < 64 || tu.block > 256 || tu.block % 64 != 0) {
    throw std::invalid_argument("launch tuning: block must be a multiple of 64 in [64, 256]");
  }
  if (tu.walk_budget < 1 || tu.walk_budget > 1024) throw std::invalid_argument("launch tuning: walk_budget in [1, 1024]");
  if (tu.search_batch < 1 || tu.search_batch > 64) throw std::invalid_argument("launch tuning: search_batch in [1, 64]");
  if (tu.chunk < 0 || tu.chunk > 64) throw std::invalid_argument("launch tuning: chunk in [0, 64]");
  if (tu.head_lanes < 1 || tu.head_lanes > 64) throw std::invalid_argument("launch tuning: head_lanes in [1, 64]");
  if (tu.head_per_mille < 0 || tu.head_per_mille > 1000) {
    throw std::invalid_argument("launch tuning: head_per_mille in [0, 1000]");
  }
  if (std::memcmp(&tu, &tu_, sizeof tu) != 0) graph_valid_ = false;
  tu_ = tu;
}

void LaunchChain::set_graph(bool on) {
  graph_on_ = on;
  if (!on) drop_graph();
}

ChainArgs LaunchChain::args(const WalkTuning& tu) const {
  ChainArgs a{};
  a.cams = d_cams_.as<const CameraDev>();
  a.n_cams = n_cams_;
  a.n_total = n_total_;
  a.groups = groups_;
  a.instant = instant_;
  a.noise = noise_;
  a.call = d_call_.as<const CallDev>();
  a.seg_capacity = seg_capacity_;
  a.run_capacity = run_capacity_;
  a.key = d_key_.as<uint8_t>();
  a.hist = d_hist_.as<uint32_t>();
  a.perm = d_perm_.as<uint32_t>();
  a.queue = d_queue_.as<unsigned int>();
  a.tu = tu;
  // The head's groups: its share of the order in groups of head_lanes, never past the pixel space.
  const uint64_t lanes = static_cast<uint64_t>(tu.head_lanes);
  const uint64_t head = static_cast<uint64_t>(n_total_) * static_cast<uint64_t>(tu.head_per_mille) / 1000;
  a.tu.head_groups = static_cast<int32_t>(std::min<uint64_t>((head + lanes - 1) / lanes, n_total_ / lanes));
  return a;
}

void LaunchChain::upload(const std::vector<CameraDev>& cams, int64_t t_end, hipStream_t stream) {
  // The staging is rewritten only once the previous upload from it has been read.
  hip_check(hipEventSynchronize(staged_), "staging wait");
  const int n = static_cast<int>(cams.size());
  if (n > h_cap_) {
    if (h_cams_ != nullptr) hip_check(hipHostFree(h_cams_), "free staging");
    h_cams_ = nullptr;
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_cams_), n * sizeof(CameraDev), hipHostMallocDefault),
              "staging");
    h_cap_ = n;
  }
  if (h_call_ == nullptr)
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_call_), sizeof(CallDev), hipHostMallocDefault), "staging");
  std::memcpy(h_cams_, cams.data(), n * sizeof(CameraDev));
  h_call_->t_end = t_end;
  int64_t rb = 0, sb = 0;
  for (const CameraDev& c : cams) {
    rb = std::max(rb, c.run_base + round_up(c.n_runs, 256));
    sb = std::max(sb, c.seg_base + round_up(c.total, kSegmentAlign));
  }
  h_call_->n_runs = rb;
  h_call_->n_segments = sb;
  hip_check(hipMemcpyAsync(d_cams_.get(), h_cams_, n * sizeof(CameraDev), hipMemcpyHostToDevice, stream),
            "upload cameras");
  hip_check(hipMemcpyAsync(d_call_.get(), h_call_, sizeof(CallDev), hipMemcpyHostToDevice, stream), "upload call");
  hip_check(hipEventRecord(staged_, stream), "event record");
}

void LaunchChain::capture() {
  drop_graph();
  if (capture_stream_ == nullptr)
    hip_check(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking), "capture stream");
  const ChainArgs a = args(tu_);
  (void)chain_walk_blocks(a);  // the occupancy query outside the capture
  hip_check(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal), "begin capture");
  hipError_t e = launch_chain_prepare(a, capture_stream_);
  if (e == hipSuccess) e = launch_chain_walk(a, capture_stream_);
  const hipError_t e_end = hipStreamEndCapture(capture_stream_, &graph_);
  hip_check(e, "capture launch chain");
  hip_check(e_end, "end capture");
  hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "instantiate graph");
  graph_valid_ = true;
  ++stats_.captures;
}

void LaunchChain::launch(std::vector<CameraDev>& cams, bool instant, int64_t t_end, hipStream_t stream) {
  if (cams.empty()) throw std::invalid_argument("launch chain: no cameras");
  uint64_t base = 0;
  int64_t rb = 0, sb = 0;
  int groups = 1;
  for (CameraDev& c : cams) {
    c.base = static_cast<uint32_t>(base);
    c.run_base = rb;
    c.seg_base = sb;
    base += static_cast<uint64_t>(c.kc.n_pixels);
    rb += round_up(c.n_runs, 256);
    sb += round_up(c.total, kSegmentAlign);
    groups = std::max(groups, c.kc.groups);
  }
  if (base > 0xffffffffull) throw std::invalid_argument("launch chain: more than 2^32 - 1 pixels in one launch");
  const bool noise = !instant && cams[0].kc.noise.tab != nullptr;
  for (const CameraDev& c : cams) {
    if ((!instant && c.kc.noise.tab != nullptr) != noise) {
      throw std::invalid_argument("launch chain: the cameras of one launch must all have the in-path noise on, or all off");
    }
  }
  const int n = static_cast<int>(cams.size());
  if (n != n_cams_ || base != n_total_ || groups != groups_ || instant != instant_ || noise != noise_) {
    graph_valid_ = false;
  }
  n_cams_ = n;
  n_total_ = static_cast<uint32_t>(base);
  groups_ = groups;
  instant_ = instant;
  noise_ = noise;
  if (!graph_on_) {
    run_capacity_ = rb;
    seg_capacity_ = sb;
  } else if (rb > run_capacity_ || sb > seg_capacity_) {
    // grow with headroom, so that the calls that follow (about the same size) replay the graph
    run_capacity_ = std::max(run_capacity_, round_up(rb + rb / 4, 256));
    seg_capacity_ = std::max(seg_capacity_, round_up(sb + sb / 4, kSegmentAlign));
    graph_valid_ = false;
  }
  const void* before[6] = {d_cams_.get(), d_call_.get(), d_key_.get(), d_perm_.get(), d_hist_.get(), d_queue_.get()};
  d_cams_.ensure(n * sizeof(CameraDev));
  d_call_.ensure(sizeof(CallDev));
  d_key_.ensure(std::max<size_t>(n_total_, 1));
  d_perm_.ensure(std::max<size_t>(n_total_, 1) * sizeof(uint32_t));
  d_hist_.ensure(128 * sizeof(uint32_t));
  d_queue_.ensure(64);
  const void* after[6] = {d_cams_.get(), d_call_.get(), d_key_.get(), d_perm_.get(), d_hist_.get(), d_queue_.get()};
  if (std::memcmp(before, after, sizeof before) != 0) graph_valid_ = false;
  t_end_ = t_end;
  upload(cams, t_end, stream);
  ++stats_.launches;
  if (graph_on_) {
    if (!graph_valid_) capture();
    hip_check(hipGraphLaunch(exec_, stream), "graph launch");
    ++stats_.graph_launches;
  } else {
    const ChainArgs a = args(tu_);
    hip_check(launch_chain_prepare(a, stream), "launch chain");
    hip_check(launch_chain_walk(a, stream), "walk");
  }
}

void LaunchChain::rerun_walk(const std::vector<CameraDev>& cams, hipStream_t stream) {
  if (static_cast<int>(cams.size()) != n_cams_) throw std::logic_error("launch chain: re-run with other cameras");
  upload(cams, t_end_, stream);
  const ChainArgs a = args(tu_);
  hip_check(launch_chain_reset(a, stream), "reset counters");
  hip_check(launch_chain_walk(a, stream), "walk");
  ++stats_.walk_reruns;
}

void LaunchChain::walk_with(const WalkTuning& tu, hipStream_t stream) {
  if (n_cams_ == 0) throw std::logic_error("launch chain: no launch to re-run");
  const ChainArgs a = args(tu);
  hip_check(launch_chain_reset(a, stream), "reset counters");
  hip_check(launch_chain_walk(a, stream), "walk");
}

}  // namespace bsx
