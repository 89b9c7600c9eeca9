// The event stream's host side: chunk slots, the copy stream and the delivery thread.
#include "stream.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>

#include "output_kernels.h"

namespace bsx {

namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

EventStream::EventStream(size_t chunk_bytes, Sink sink, void* user) : sink_(sink), user_(user) {
  if (sink == nullptr) throw std::invalid_argument("EventStream: null sink");
  if (chunk_bytes < sizeof(StreamEventDev) || chunk_bytes > kMaxChunkBytes || chunk_bytes % sizeof(StreamEventDev)) {
    throw std::invalid_argument("EventStream: chunk_bytes must be a multiple of 16 in [16, 2^32]");
  }
  chunk_events_ = chunk_bytes / sizeof(StreamEventDev);
  try {
    for (int s = 0; s < 2; ++s) {
      hip_check(hipMalloc(reinterpret_cast<void**>(&dev_[s]), chunk_bytes), "event stream device chunk");
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_[s]), chunk_bytes, hipHostMallocDefault),
                "event stream pinned chunk");
      hip_check(hipEventCreateWithFlags(&packed_[s], hipEventDisableTiming), "hipEventCreate");
      hip_check(hipEventCreate(&copy_start_[s]), "hipEventCreate");
      hip_check(hipEventCreate(&copied_[s]), "hipEventCreate");
    }
    hip_check(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "event stream copy stream");
  } catch (...) {
    for (int s = 0; s < 2; ++s) {
      if (dev_[s] != nullptr) (void)hipFree(dev_[s]);
      if (host_[s] != nullptr) (void)hipHostFree(host_[s]);
      if (packed_[s] != nullptr) (void)hipEventDestroy(packed_[s]);
      if (copy_start_[s] != nullptr) (void)hipEventDestroy(copy_start_[s]);
      if (copied_[s] != nullptr) (void)hipEventDestroy(copied_[s]);
    }
    if (copy_ != nullptr) (void)hipStreamDestroy(copy_);
    throw;
  }
  thread_ = std::thread([this] { deliver_loop(); });
}

EventStream::~EventStream() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
    drop_ = true;
  }
  cv_.notify_all();
  thread_.join();
  (void)hipStreamSynchronize(copy_);
  for (int s = 0; s < 2; ++s) {
    (void)hipFree(dev_[s]);
    (void)hipHostFree(host_[s]);
    (void)hipEventDestroy(packed_[s]);
    (void)hipEventDestroy(copy_start_[s]);
    (void)hipEventDestroy(copied_[s]);
  }
  (void)hipStreamDestroy(copy_);
}

void EventStream::raise_if_failed() {
  std::lock_guard<std::mutex> lk(mu_);
  if (!error_.empty()) throw std::runtime_error(error_);
}

void EventStream::wait_slot(int slot) {
  const auto t0 = std::chrono::steady_clock::now();
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return !held_[slot] || !error_.empty(); });
  stats_.wait_ms += ms_since(t0);
  if (!error_.empty()) throw std::runtime_error(error_);
}

void EventStream::push(const OutEventDev* ev, size_t n, hipStream_t stream) {
  raise_if_failed();
  if (n > 0 && ev == nullptr) throw std::invalid_argument("EventStream.push: null events");
  size_t done = 0;
  while (done < n) {
    if (fill_ == 0) wait_slot(slot_);  // the slot's last copy and delivery are over
    const size_t take = std::min(n - done, chunk_events_ - fill_);
    hip_check(launch_out_pack(ev + done, take, dev_[slot_] + fill_, stream), "pack events");
    fill_ += take;
    done += take;
    {
      std::lock_guard<std::mutex> lk(mu_);
      stats_.events += take;
    }
    if (fill_ == chunk_events_) submit(stream);
  }
}

void EventStream::submit(hipStream_t stream) {
  const int s = slot_;
  hip_check(hipEventRecord(packed_[s], stream), "event record");
  hip_check(hipStreamWaitEvent(copy_, packed_[s], 0), "copy stream wait");
  hip_check(hipEventRecord(copy_start_[s], copy_), "event record");
  hip_check(hipMemcpyAsync(host_[s], dev_[s], fill_ * sizeof(StreamEventDev), hipMemcpyDeviceToHost, copy_),
            "chunk copy");
  hip_check(hipEventRecord(copied_[s], copy_), "event record");
  {
    std::lock_guard<std::mutex> lk(mu_);
    held_[s] = true;
    jobs_.push_back(Job{s, fill_, next_index_++});
  }
  cv_.notify_all();
  slot_ = 1 - s;
  fill_ = 0;
}

void EventStream::finish(hipStream_t stream) {
  raise_if_failed();
  if (fill_ > 0) submit(stream);
  const auto t0 = std::chrono::steady_clock::now();
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return jobs_.empty() || !error_.empty(); });
  stats_.wait_ms += ms_since(t0);
  if (!error_.empty()) throw std::runtime_error(error_);
}

EventStreamStats EventStream::stats() const {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

void EventStream::deliver_loop() {
  while (true) {
    Job j{};
    bool failed = false;  // after a failure the remaining chunks are released undelivered
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
      if (jobs_.empty() || drop_) return;
      j = jobs_.front();
      failed = !error_.empty();
    }
    std::string failure;
    float ms = 0.0f;
    hipError_t e = hipEventSynchronize(copied_[j.slot]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, copy_start_[j.slot], copied_[j.slot]);
    if (e != hipSuccess) {
      failure = std::string("EventStream: chunk copy: ") + hipGetErrorString(e);
    } else if (!failed) {
      const int rc = sink_(host_[j.slot], j.count, j.index, user_);
      if (rc != 0) failure = "EventStream: the sink returned " + std::to_string(rc) + " for chunk " + std::to_string(j.index);
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_.pop_front();
      held_[j.slot] = false;
      if (failure.empty() && !failed) {
        ++stats_.chunks;
        stats_.bytes += j.count * sizeof(StreamEventDev);
        stats_.copy_ms += ms;
      } else if (error_.empty()) {
        error_ = failure;
      }
    }
    cv_.notify_all();
  }
}

}  // namespace bsx
