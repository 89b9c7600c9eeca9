/* bsx output stage: a camera's events in output order, window by window, as C4 columns.
 *
 * Reference: the engine's per-camera output stage, Camera::output (src/engine/engine_impl.h:1761-1840)
 * with CropFlags / BuildEvents / OutKeys / GatherEvents (src/engine/output.h:29-120), the order keys,
 * S1 labels and window cut of src/sensor/output_order.h:22-120, the readout (src/sensor/readout.h:22-26)
 * and the C4 event columns of python/ballasim/engine/writer.py:1-40. The ground-truth columns (centre-ray
 * instance, depth, flow) need the renderer and are not produced.
 *
 * Per window: the window's raw device events (bsx_sensor_advance_device's, or a rig's) inside the crop
 * get their reported time and S1 label, join the events held from earlier windows, are ordered by
 * (t_reported, x, y, t_cross) on the device, and the events reported before the window's end on the
 * sensor clock (offset and drift, no delay or quantum) are released; the rest stay held. Conventions
 * as bsx_sensor.h.
 */
#ifndef BSX_OUTPUT_H_
#define BSX_OUTPUT_H_

#include "bsx_sensor.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bsx_output bsx_output;

typedef struct bsx_crop {
  int32_t x0, y0, x1, y1; /* kept: x0 <= x < x1, y0 <= y < y1 */
} bsx_crop;

/* Released events (host arrays owned by the stage, valid until its next call). C4 columns: t in us on
 * the sensor clock relative to t_offset_us, t_ns_rem = t_reported - 1000 floor(t_reported / 1000);
 * label in S1 classes (0 signal, 2 noise, 3 leak, 4 hot), flags the sensor's bits. */
typedef struct bsx_output_events {
  size_t count;
  const uint16_t* x;
  const uint16_t* y;
  const uint8_t* p;
  const int64_t* t;
  const uint16_t* t_ns_rem;
  const int64_t* t_reported_ns;
  const int64_t* t_cross_ns;
  const uint8_t* label;
  const uint16_t* flags;
  size_t held; /* events held for later windows after this call */
} bsx_output_events;

/* The stage of one camera (its size and readout are read from the model, which must outlive the
 * stage); crop null: every event kept. */
int bsx_output_create(bsx_sensor* camera, const bsx_crop* crop, int64_t t_offset_us, bsx_output** out);
int bsx_output_destroy(bsx_output* o);
/* Window [t_w, t_w + 2^23 ns): its raw device events in, the released ones out. */
int bsx_output_window(bsx_output* o, const bsx_device_events* events, int64_t t_w, void* stream,
                      bsx_output_events* out);
/* Every held event, in output order (the end of a run). */
int bsx_output_flush(bsx_output* o, void* stream, bsx_output_events* out);

#ifdef __cplusplus
}
#endif

#endif /* BSX_OUTPUT_H_ */
