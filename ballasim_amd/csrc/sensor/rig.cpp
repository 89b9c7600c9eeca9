// A rig of cameras advanced by one launch chain, and its autotuner.
#include "rig.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace bsx {

namespace {

constexpr int kK = 1 + kStCount;  // counter words per camera

bool same_tuning(const WalkTuning& a, const WalkTuning& b) {
  return a.block == b.block && a.walk_budget == b.walk_budget && a.search_batch == b.search_batch &&
         a.chunk == b.chunk && a.head_lanes == b.head_lanes && a.head_per_mille == b.head_per_mille;
}

}  // namespace

Rig::Rig(std::vector<SensorModel*> cameras) : cams_(std::move(cameras)) {
  if (cams_.empty() || cams_.size() > 64) throw std::invalid_argument("Rig: need 1 to 64 cameras");
  for (size_t c = 0; c < cams_.size(); ++c) {
    if (cams_[c] == nullptr) throw std::invalid_argument("Rig: null camera");
    for (size_t d = 0; d < c; ++d)
      if (cams_[d] == cams_[c]) throw std::invalid_argument("Rig: a camera is given twice");
    if (cams_[c]->instant() != cams_[0]->instant()) {
      throw std::invalid_argument("Rig: all cameras must share one photoreceptor kind (instantaneous or cascade)");
    }
  }
  counters_.assign(cams_.size() * kK, 0ull);
  hip_check(hipEventCreate(&ev0_), "hipEventCreate");
  hip_check(hipEventCreate(&ev1_), "hipEventCreate");
}

Rig::~Rig() {
  if (ev0_ != nullptr) (void)hipEventDestroy(ev0_);
  if (ev1_ != nullptr) (void)hipEventDestroy(ev1_);
}

RunsDev Rig::device_runs(size_t c, const bsx_device_runs& r) const {
  RunsDev d{};
  d.row_of = cams_[c]->d_row_of_.as<const int32_t>();
  d.row_start = r.row_start;
  d.row_count = r.row_count;
  d.t = r.t;
  d.irr = r.irr;
  return d;
}

std::vector<CameraDev> Rig::check_and_stage(const bsx_device_runs* runs, int64_t t_end_ns, hipStream_t stream) {
  if (runs == nullptr) throw std::invalid_argument("Rig: null runs");
  for (size_t c = 0; c < cams_.size(); ++c) {
    try {
      cams_[c]->enqueue_device_checks(runs[c], t_end_ns, stream);
    } catch (const std::invalid_argument& e) {
      throw std::invalid_argument("Rig camera " + std::to_string(c) + ": " + e.what());
    } catch (const std::logic_error& e) {
      throw std::logic_error("Rig camera " + std::to_string(c) + ": " + e.what());
    }
  }
  hip_check(hipStreamSynchronize(stream), "rig checks");
  for (size_t c = 0; c < cams_.size(); ++c) {
    try {
      cams_[c]->finish_device_checks();
    } catch (const std::invalid_argument& e) {
      throw std::invalid_argument("Rig camera " + std::to_string(c) + ": " + e.what());
    }
  }
  std::vector<CameraDev> cds;
  cds.reserve(cams_.size());
  for (size_t c = 0; c < cams_.size(); ++c) {
    SensorModel& m = *cams_[c];
    cds.push_back(m.stage(device_runs(c, runs[c]), runs[c].pixel, runs[c].n, m.scan_total_));
  }
  return cds;
}

void Rig::refuse_non_finite() {
  for (size_t c = 0; c < cams_.size(); ++c) {
    try {
      cams_[c]->refuse_non_finite(counters_.data() + c * kK);
    } catch (const std::invalid_argument& e) {
      for (SensorModel* m : cams_) m->initialized_ = true;  // no camera advanced
      throw std::invalid_argument("Rig camera " + std::to_string(c) + ": " + e.what());
    }
  }
}

void Rig::advance_device(const bsx_device_runs* runs, int64_t t_end_ns, hipStream_t stream, bsx_device_events* out) {
  if (out == nullptr) throw std::invalid_argument("Rig: null events output");
  std::vector<CameraDev> cds = check_and_stage(runs, t_end_ns, stream);
  for (SensorModel* m : cams_) m->initialized_ = false;
  const bool instant = cams_[0]->instant();
  hip_check(hipEventRecord(ev0_, stream), "event record");
  chain_.launch(cds, instant, t_end_ns, stream);
  while (true) {
    hip_check(hipEventRecord(ev1_, stream), "event record");
    for (size_t c = 0; c < cams_.size(); ++c) {
      hip_check(hipMemcpyAsync(counters_.data() + c * kK, cams_[c]->d_counters_.get(), kK * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, stream),
                "read counters");
    }
    hip_check(hipStreamSynchronize(stream), "rig advance");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event time");
    kernel_ms_ += ms;
    refuse_non_finite();
    bool overflow = false;
    for (size_t c = 0; c < cams_.size(); ++c) {
      const unsigned long long n = counters_[c * kK];
      if (n > cds[c].cap) {
        cams_[c]->grow_events(n, cds[c]);
        overflow = true;
      }
    }
    if (!overflow) break;
    hip_check(hipEventRecord(ev0_, stream), "event record");
    chain_.rerun_walk(cds, stream);
  }
  for (size_t c = 0; c < cams_.size(); ++c) {
    SensorModel& m = *cams_[c];
    m.commit(counters_.data() + c * kK, t_end_ns);
    m.stats_.input_clamps += counters_[c * kK + 1 + kStClamps];
    ++m.stats_.launches;
    m.initialized_ = true;
    out[c] = bsx_device_events{reinterpret_cast<const bsx_raw_event*>(m.d_events_.get()),
                               static_cast<size_t>(counters_[c * kK])};
  }
}

std::vector<RigTrial> Rig::autotune(const bsx_device_runs* runs, int64_t t_end_ns, hipStream_t stream, int reps) {
  if (reps < 1 || reps > 100) throw std::invalid_argument("Rig.autotune: reps in [1, 100]");
  std::vector<CameraDev> cds = check_and_stage(runs, t_end_ns, stream);
  const bool instant = cams_[0]->instant();
  chain_.launch(cds, instant, t_end_ns, stream);  // phase A and the order, reused by every trial
  for (size_t c = 0; c < cams_.size(); ++c) {
    hip_check(hipMemcpyAsync(counters_.data() + c * kK, cams_[c]->d_counters_.get(), kK * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, stream),
              "read counters");
  }
  hip_check(hipStreamSynchronize(stream), "rig autotune");
  refuse_non_finite();
  std::vector<RigTrial> trials;
  const auto time_walk = [&](const WalkTuning& tu) {
    for (const RigTrial& r : trials)
      if (same_tuning(r.tuning, tu)) return r.walk_ms;
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
      hip_check(hipEventRecord(ev0_, stream), "event record");
      chain_.walk_with(tu, stream);
      hip_check(hipEventRecord(ev1_, stream), "event record");
      hip_check(hipStreamSynchronize(stream), "rig autotune");
      float ms = 0.0f;
      hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event time");
      best = std::min(best, static_cast<double>(ms));
    }
    trials.push_back(RigTrial{tu, best});
    return best;
  };
  // Coordinate descent: each axis in turn, every value, the best kept; two passes.
  static const int kBlocks[] = {64, 128, 256};
  static const int kBudgets[] = {4, 8, 16, 32, 64};
  static const int kBatches[] = {1, 4, 8, 16, 32, 64};
  static const int kChunks[] = {0, 1, 2, 4, 8, 16};
  // The head (lanes, thousandths of the order): none, or the heaviest pixels in groups of 16 or 32.
  static const int kHeads[][2] = {{64, 0}, {16, 8}, {32, 8}, {32, 32}};
  WalkTuning best = chain_.tuning();
  double best_ms = time_walk(best);
  for (int pass = 0; pass < 2; ++pass) {
    for (int axis = 0; axis < 5; ++axis) {
      const int* vals = axis == 0 ? kBlocks : axis == 1 ? kBudgets : axis == 2 ? kBatches : kChunks;
      const int nv = axis == 0 ? 3 : axis == 1 ? 5 : axis == 4 ? 4 : 6;
      const WalkTuning base = best;
      for (int k = 0; k < nv; ++k) {
        WalkTuning t = base;
        if (axis == 4) {
          t.head_lanes = kHeads[k][0];
          t.head_per_mille = kHeads[k][1];
        } else {
          (axis == 0 ? t.block : axis == 1 ? t.walk_budget : axis == 2 ? t.search_batch : t.chunk) = vals[k];
        }
        const double ms = time_walk(t);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
    }
  }
  chain_.set_tuning(best);
  return trials;
}

}  // namespace bsx
