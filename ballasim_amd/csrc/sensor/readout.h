// The readout of a threshold crossing on the host and the device (reference: Readout and
// reported_time, src/sensor/readout.h:13-26; output_cut, src/sensor/output_order.h:50-56): the report
// is delayed by delay_ns, floored to the counter quantum (phase per camera) and expressed on the
// sensor clock t + offset + llround(drift t).
#pragma once

#include <math.h>
#include <stdint.h>

#include "../core/hd.h"

namespace bsx {

struct Readout {
  int64_t quantum_ns, delay_ns, phase_ns, clock_offset_ns;
  double clock_drift_ppm;
};

BSX_HD BSX_INLINE int64_t floor_div64(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

BSX_HD BSX_INLINE int64_t reported_time(const Readout& r, int64_t t_cross_ns) {
  int64_t t = t_cross_ns + r.delay_ns;
  if (r.quantum_ns > 0) t = r.phase_ns + r.quantum_ns * floor_div64(t - r.phase_ns, r.quantum_ns);
  return t + r.clock_offset_ns + static_cast<int64_t>(llround(r.clock_drift_ppm * 1e-6 * static_cast<double>(t)));
}

// The window cut: events reported before the end of window t_end on the sensor clock are final (the
// readout with zero delay and no quantum).
BSX_HD BSX_INLINE int64_t output_cut(const Readout& r, int64_t t_window_end_ns) {
  Readout c = r;
  c.delay_ns = 0;
  c.quantum_ns = 0;
  return reported_time(c, t_window_end_ns);
}

}  // namespace bsx
