// Released events streamed to the host in double-buffered chunks of pinned memory.
//
// Events pushed in output order are packed on the device (16-byte StreamEventDev records) into the
// current one of two device chunks. A full chunk (or the partial last one at finish) is copied to its
// pinned host buffer by an async copy on the stream's own copy stream, which waits for the packing
// through an event; a delivery thread waits for that copy and hands the host chunk to the sink. The
// two chunk slots alternate, so the packing of one chunk (and the compute queued around it) overlaps
// the copy and the delivery of the other. A slot is reused once the sink has returned for it: a push
// that needs a slot still held blocks until then (the producer waits on the consumer, never the other
// way round).
//
// The reference has no counterpart: its engine hands each window's events to the writer on the host
// (python/ballasim/engine/writer.py:1-40). This is the stress configuration's event path (a 2048x1536
// sensor over 1 s, about 1e9 events), where the events cannot be gathered in one piece.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>

#include "output.h"

namespace bsx {

struct EventStreamStats {
  uint64_t chunks = 0;  // chunks delivered to the sink
  uint64_t events = 0;  // events pushed
  uint64_t bytes = 0;   // bytes copied to the host
  double wait_ms = 0;   // time push and finish spent waiting for a chunk slot or a delivery
  double copy_ms = 0;   // device-to-host copy time of the delivered chunks (HIP events on the copy stream)
};

class EventStream {
 public:
  // sink(events, count, index, user): chunk `index` (0, 1, ...) of count events in pinned host memory,
  // valid until the sink returns; a nonzero return stops the stream (later calls throw).
  using Sink = int (*)(const StreamEventDev* events, size_t count, uint64_t index, void* user);
  static constexpr size_t kMaxChunkBytes = size_t{1} << 32;

  // chunk_bytes: a multiple of 16 in [16, 2^32].
  EventStream(size_t chunk_bytes, Sink sink, void* user);
  ~EventStream();  // chunks not yet delivered are dropped (call finish first)
  EventStream(const EventStream&) = delete;
  EventStream& operator=(const EventStream&) = delete;

  // Queues the packing of ev[0 .. n) (device memory, output order) on stream, after the work already
  // queued there that produces them; ev may be reused once the stream has passed this point.
  void push(const OutEventDev* ev, size_t n, hipStream_t stream);
  // Sends the partial chunk and returns once every chunk has been delivered.
  void finish(hipStream_t stream);
  EventStreamStats stats() const;
  size_t chunk_events() const { return chunk_events_; }

 private:
  struct Job {
    int slot;
    size_t count;
    uint64_t index;
  };
  void submit(hipStream_t stream);
  void wait_slot(int slot);
  void deliver_loop();
  void raise_if_failed();

  size_t chunk_events_;
  Sink sink_;
  void* user_;
  StreamEventDev* dev_[2] = {nullptr, nullptr};
  StreamEventDev* host_[2] = {nullptr, nullptr};
  hipStream_t copy_ = nullptr;
  hipEvent_t packed_[2] = {nullptr, nullptr}, copy_start_[2] = {nullptr, nullptr}, copied_[2] = {nullptr, nullptr};
  int slot_ = 0;     // the slot being filled
  size_t fill_ = 0;  // events in it
  uint64_t next_index_ = 0;

  mutable std::mutex mu_;
  std::condition_variable cv_;
  bool held_[2] = {false, false};  // submitted and not yet delivered
  std::deque<Job> jobs_;
  bool stop_ = false, drop_ = false;
  std::string error_;
  EventStreamStats stats_;
  std::thread thread_;
};

}  // namespace bsx
