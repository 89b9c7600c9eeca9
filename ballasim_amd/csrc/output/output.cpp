// The output stage's host side: window and flush orchestration over output.hip's kernels.
#include "output.h"

#include <algorithm>
#include <stdexcept>

#include "output_kernels.h"

namespace bsx {

namespace {

int bit_length(uint64_t v) {
  int b = 0;
  while (v != 0) {
    ++b;
    v >>= 1;
  }
  return b;
}

}  // namespace

void OutputColumns::resize(size_t n) {
  x.resize(n);
  y.resize(n);
  p.resize(n);
  t_us.resize(n);
  t_ns_rem.resize(n);
  t_reported.resize(n);
  t_cross.resize(n);
  label.resize(n);
  flags.resize(n);
}

OutputStage::OutputStage(uint32_t width, uint32_t height, const Readout& readout, const CropBox* crop,
                         int64_t t_offset_us)
    : width_(width), height_(height), readout_(readout), crop_on_(crop != nullptr), t_offset_us_(t_offset_us) {
  if (width == 0 || height == 0) throw std::invalid_argument("OutputStage: empty camera");
  if (crop != nullptr) {
    if (!(crop->x0 >= 0 && crop->y0 >= 0 && crop->x0 < crop->x1 && crop->y0 < crop->y1 &&
          crop->x1 <= static_cast<int32_t>(width) && crop->y1 <= static_cast<int32_t>(height))) {
      throw std::invalid_argument("OutputStage: the crop must lie inside the frame and be non-empty");
    }
    crop_ = *crop;
  }
  y_bits_ = 1;
  while ((uint64_t{1} << y_bits_) < height) ++y_bits_;
  scratch_.ensure(kOutScratch * 8);
}

size_t OutputStage::take_window(const RawEventDev* raw, size_t n, int64_t t_w, hipStream_t stream, int64_t* cut) {
  if (n > 0 && raw == nullptr) throw std::invalid_argument("OutputStage.window: null events");
  const size_t n_max = n_held_ + n;
  if (n_max > 0xffffffffull) throw std::invalid_argument("OutputStage.window: at most 2^32 - 1 events");
  // The held events stay at the front of held_; the window's kept events are appended after them.
  if (held_.bytes() < std::max<size_t>(n_max, 1) * sizeof(OutEventDev)) {
    DeviceBuffer grown;
    grown.ensure(std::max<size_t>(n_max + n_max / 2, 1024) * sizeof(OutEventDev));
    if (n_held_ > 0) {
      hip_check(hipMemcpyAsync(grown.get(), held_.get(), n_held_ * sizeof(OutEventDev), hipMemcpyDeviceToDevice,
                               stream),
                "grow held");
    }
    hip_check(hipStreamSynchronize(stream), "grow held");
    held_.swap(grown);
  }
  unsigned long long init[kOutScratch] = {0, ~0ull, 0, ~0ull, 0, 0, 0, 0};
  hip_check(hipMemcpyAsync(scratch_.get(), init, sizeof(init), hipMemcpyHostToDevice, stream), "scratch");
  unsigned long long* sc = scratch_.as<unsigned long long>();
  hip_check(launch_out_build(raw, n, width_, readout_, crop_on_, crop_, held_.as<OutEventDev>() + n_held_, sc, stream),
            "build events");
  hip_check(hipStreamSynchronize(stream), "build events");  // init[] must outlive its copy
  *cut = output_cut(readout_, t_w + kPlainSegmentNs);
  hip_check(launch_out_minmax(held_.as<const OutEventDev>(), n_held_, n_max, sc, stream), "min/max");
  unsigned long long h[kOutScratch];
  hip_check(hipMemcpyAsync(h, scratch_.get(), sizeof(h), hipMemcpyDeviceToHost, stream), "read scratch");
  hip_check(hipStreamSynchronize(stream), "output window");
  return n_held_ + static_cast<size_t>(h[0]);
}

size_t OutputStage::take_flush(hipStream_t stream) {
  unsigned long long init[kOutScratch] = {0, ~0ull, 0, ~0ull, 0, 0, 0, 0};
  hip_check(hipMemcpyAsync(scratch_.get(), init, sizeof(init), hipMemcpyHostToDevice, stream), "scratch");
  hip_check(launch_out_minmax(held_.as<const OutEventDev>(), n_held_, n_held_, scratch_.as<unsigned long long>(),
                              stream),
            "min/max");
  hip_check(hipStreamSynchronize(stream), "output flush");
  return n_held_;
}

const OutputColumns& OutputStage::window(const RawEventDev* raw, size_t n, int64_t t_w, hipStream_t stream) {
  int64_t cut = 0;
  const size_t total = take_window(raw, n, t_w, stream, &cut);
  return download(release_sorted(total, cut, stream), stream);
}

const OutputColumns& OutputStage::flush(hipStream_t stream) {
  const size_t total = take_flush(stream);
  return download(release_sorted(total, INT64_MAX, stream), stream);
}

size_t OutputStage::window_device(const RawEventDev* raw, size_t n, int64_t t_w, hipStream_t stream,
                                  const OutEventDev** released) {
  int64_t cut = 0;
  const size_t total = take_window(raw, n, t_w, stream, &cut);
  const size_t m = release_sorted(total, cut, stream);
  *released = sorted_.as<const OutEventDev>();
  return m;
}

size_t OutputStage::flush_device(hipStream_t stream, const OutEventDev** released) {
  const size_t m = release_sorted(take_flush(stream), INT64_MAX, stream);
  *released = sorted_.as<const OutEventDev>();
  return m;
}

size_t OutputStage::release_sorted(size_t total, int64_t cut, hipStream_t stream) {
  if (total == 0) {
    n_held_ = 0;
    return 0;
  }
  unsigned long long h[kOutScratch];
  hip_check(hipMemcpyAsync(h, scratch_.get(), sizeof(h), hipMemcpyDeviceToHost, stream), "read scratch");
  hip_check(hipStreamSynchronize(stream), "output min/max");
  const uint64_t bias = uint64_t{1} << 63;
  const int64_t tc_min = static_cast<int64_t>(h[1] ^ bias), tc_max = static_cast<int64_t>(h[2] ^ bias);
  const int64_t tr_min = static_cast<int64_t>(h[3] ^ bias), tr_max = static_cast<int64_t>(h[4] ^ bias);
  keys_.ensure(total * 8);
  keys_alt_.ensure(total * 8);
  perm_.ensure(total * 4);
  perm_alt_.ensure(total * 4);
  sorted_.ensure(total * sizeof(OutEventDev));
  const OutEventDev* ev = held_.as<const OutEventDev>();
  // Three stable passes, least significant first: t_cross, (x, y), t_reported.
  const int bits[3] = {std::max(1, bit_length(static_cast<uint64_t>(tc_max - tc_min))), 16 + y_bits_,
                       std::max(1, bit_length(static_cast<uint64_t>(tr_max - tr_min)))};
  for (int pass = 0; pass < 3; ++pass) {
    hip_check(launch_out_keys(ev, pass == 0 ? nullptr : perm_.as<const uint32_t>(),
                              pass == 0 ? perm_.as<uint32_t>() : nullptr, total, pass, y_bits_, tc_min, tr_min,
                              keys_.as<uint64_t>(), stream),
              "sort keys");
    hip_check(launch_out_sort(keys_.as<uint64_t>(), keys_alt_.as<uint64_t>(), perm_.as<uint32_t>(),
                              perm_alt_.as<uint32_t>(), total, bits[pass], temp_, stream),
              "radix sort");
    perm_.swap(perm_alt_);
  }
  hip_check(launch_out_gather(ev, perm_.as<const uint32_t>(), total, cut, sorted_.as<OutEventDev>(),
                              scratch_.as<unsigned long long>(), stream),
            "gather");
  unsigned long long ready = 0;
  hip_check(hipMemcpyAsync(&ready, scratch_.as<unsigned long long>() + 5, 8, hipMemcpyDeviceToHost, stream),
            "read released");
  hip_check(hipStreamSynchronize(stream), "output gather");
  const size_t m = static_cast<size_t>(ready);
  // The unreleased rest, in output order, becomes the held buffer.
  n_held_ = total - m;
  if (n_held_ > 0) {
    hip_check(hipMemcpyAsync(held_.get(), sorted_.as<OutEventDev>() + m, n_held_ * sizeof(OutEventDev),
                             hipMemcpyDeviceToDevice, stream),
              "hold");
  }
  return m;
}

const OutputColumns& OutputStage::download(size_t m, hipStream_t stream) {
  out_.resize(m);
  if (m == 0) return out_;
  cx_.ensure(m * 2);
  cy_.ensure(m * 2);
  cp_.ensure(m);
  ct_.ensure(m * 8);
  crem_.ensure(m * 2);
  ctr_.ensure(m * 8);
  ctc_.ensure(m * 8);
  clab_.ensure(m);
  cfl_.ensure(m * 2);
  hip_check(launch_out_columns(sorted_.as<const OutEventDev>(), m, t_offset_us_, cx_.as<uint16_t>(), cy_.as<uint16_t>(),
                               cp_.as<uint8_t>(), ct_.as<int64_t>(), crem_.as<uint16_t>(), ctr_.as<int64_t>(),
                               ctc_.as<int64_t>(), clab_.as<uint8_t>(), cfl_.as<uint16_t>(), stream),
            "columns");
  const auto get = [&](void* dst, const DeviceBuffer& src, size_t bytes) {
    hip_check(hipMemcpyAsync(dst, src.get(), bytes, hipMemcpyDeviceToHost, stream), "download columns");
  };
  get(out_.x.data(), cx_, m * 2);
  get(out_.y.data(), cy_, m * 2);
  get(out_.p.data(), cp_, m);
  get(out_.t_us.data(), ct_, m * 8);
  get(out_.t_ns_rem.data(), crem_, m * 2);
  get(out_.t_reported.data(), ctr_, m * 8);
  get(out_.t_cross.data(), ctc_, m * 8);
  get(out_.label.data(), clab_, m);
  get(out_.flags.data(), cfl_, m * 2);
  hip_check(hipStreamSynchronize(stream), "output columns");
  return out_;
}

}  // namespace bsx
