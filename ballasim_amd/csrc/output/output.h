// The output stage of a camera on the device: each window's raw events (cropped, with their reported
// times and S1 labels) join the events held from earlier windows, are put in output order
// (t_reported, x, y, t_cross) and those reported before the window's end on the sensor clock are
// released as C4 columns; the rest stay held.
//
// Reference: Camera::output (src/engine/engine_impl.h:1761-1840), CropFlags / BuildEvents / OutKeys /
// GatherEvents (src/engine/output.h:29-120), order_key / output_order / output_cut / label_s1
// (src/sensor/output_order.h:22-120), reported_time (src/sensor/readout.h:22-26) and the C4 columns
// (python/ballasim/engine/writer.py:1-40).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../sensor/kconst.h"
#include "../sensor/model.h"
#include "../sensor/readout.h"

namespace bsx {

// One event in output form (24 bytes; the reference's OutEvent without the ground-truth fields).
struct OutEventDev {
  int64_t t_reported;
  int64_t t_cross;
  uint16_t x, y, flags;
  uint8_t p, label;  // label: S1 class
};
static_assert(sizeof(OutEventDev) == 24, "OutEventDev layout");

// One released event as the event stream carries it (16 bytes): the reported time on the sensor
// clock in ns, the pixel, the polarity, the S1 label and the sensor's flags.
struct StreamEventDev {
  int64_t t_reported;
  uint16_t x, y;
  uint8_t p, label;
  uint16_t flags;
};
static_assert(sizeof(StreamEventDev) == 16, "StreamEventDev layout");

struct CropBox {
  int32_t x0, y0, x1, y1;  // kept: x0 <= x < x1, y0 <= y < y1
};

// Output columns of released events (host copies, valid until the next call).
struct OutputColumns {
  std::vector<uint16_t> x, y;
  std::vector<uint8_t> p;
  std::vector<int64_t> t_us;        // floor(t_reported / 1000) - t_offset_us
  std::vector<uint16_t> t_ns_rem;   // t_reported - 1000 floor(t_reported / 1000)
  std::vector<int64_t> t_reported, t_cross;
  std::vector<uint8_t> label;
  std::vector<uint16_t> flags;
  void resize(size_t n);
};

class OutputStage {
 public:
  // The camera's width, height and readout; crop null: every event kept.
  OutputStage(uint32_t width, uint32_t height, const Readout& readout, const CropBox* crop, int64_t t_offset_us);
  // Window [t_w, t_w + 2^23): its raw device events join the held ones; the events reported before
  // the window's end on the sensor clock are released in output order.
  const OutputColumns& window(const RawEventDev* raw, size_t n, int64_t t_w, hipStream_t stream);
  // Every held event, in output order (the end of a run).
  const OutputColumns& flush(hipStream_t stream);
  // window and flush with the released events left on the device, in output order: returns their
  // count; *released points at them until the stage's next call (work queued on stream).
  size_t window_device(const RawEventDev* raw, size_t n, int64_t t_w, hipStream_t stream,
                       const OutEventDev** released);
  size_t flush_device(hipStream_t stream, const OutEventDev** released);
  size_t held() const { return n_held_; }

 private:
  // Sorts the total events of held_, moves those reported before cut to the front of sorted_ (their
  // count returned) and the rest back to held_.
  size_t release_sorted(size_t total, int64_t cut, hipStream_t stream);
  const OutputColumns& download(size_t m, hipStream_t stream);
  // The window's raw events joined to the held ones; returns the total and sets *cut.
  size_t take_window(const RawEventDev* raw, size_t n, int64_t t_w, hipStream_t stream, int64_t* cut);
  size_t take_flush(hipStream_t stream);

  uint32_t width_, height_;
  int y_bits_;
  Readout readout_;
  bool crop_on_;
  CropBox crop_{};
  int64_t t_offset_us_;
  size_t n_held_ = 0;
  DeviceBuffer held_, sorted_, keys_, keys_alt_, perm_, perm_alt_, temp_, scratch_;
  DeviceBuffer cx_, cy_, cp_, ct_, crem_, ctr_, ctc_, clab_, cfl_;
  OutputColumns out_;
};

}  // namespace bsx
