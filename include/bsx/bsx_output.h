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

/* ---- Event stream: released events to the host in double-buffered pinned chunks ----
 *
 * No counterpart in the reference, whose engine hands each window's events to the writer on the host
 * (python/ballasim/engine/writer.py:1-40); this is the stress configuration's output path (a
 * 2048x1536 sensor over 1 s, about 1e9 events). The released events of an output stage, in output
 * order, are packed on the device into 16-byte records, gathered in two device chunks of chunk_bytes
 * that alternate, and each full chunk (the partial last one at finish) is copied to its pinned host
 * buffer by an async copy on a copy stream of the stream's own while the next chunk fills. A
 * delivery thread of the stream calls sink(chunk) once the copy has landed; the chunk's memory is
 * valid until the sink returns and is then reused. A call that needs a chunk still held by the sink
 * waits for it. Chunks arrive in order (index 0, 1, ...); concatenated they are exactly the stage's
 * released events, window after window. Device memory: 2 chunk_bytes; pinned host memory: 2
 * chunk_bytes (256 MiB chunks: 512 MiB of each). */
typedef struct bsx_event_stream bsx_event_stream;

typedef struct bsx_stream_event {
  int64_t t_reported_ns; /* reported time on the sensor clock */
  uint16_t x, y;
  uint8_t p, label;      /* polarity 0/1, S1 class */
  uint16_t flags;
} bsx_stream_event;      /* 16 bytes */

typedef struct bsx_stream_stats {
  uint64_t chunks;       /* delivered */
  uint64_t events;       /* pushed */
  uint64_t bytes;        /* copied to the host */
  double wait_ms;        /* host time spent waiting for a chunk slot or for the deliveries at finish */
  double copy_ms;        /* device-to-host copy time of the delivered chunks */
} bsx_stream_stats;

/* Called on the stream's delivery thread; a nonzero return stops the stream (later calls fail with
 * BSX_E_RUNTIME). */
typedef int (*bsx_chunk_sink)(const bsx_stream_event* events, size_t count, uint64_t index, void* user);

/* chunk_bytes: a multiple of 16 in [16, 2^32] (256 MiB for the stress configuration). */
int bsx_event_stream_create(size_t chunk_bytes, bsx_chunk_sink sink, void* user, bsx_event_stream** out);
/* Chunks not yet delivered are dropped: call bsx_event_stream_finish first. */
int bsx_event_stream_destroy(bsx_event_stream* s);
/* bsx_output_window / bsx_output_flush with the released events streamed instead of returned:
 * *released (may be null) = their count, *held (may be null) = the events held after the call. */
int bsx_output_window_to_stream(bsx_output* o, const bsx_device_events* events, int64_t t_w, void* stream,
                                bsx_event_stream* s, size_t* released, size_t* held);
int bsx_output_flush_to_stream(bsx_output* o, void* stream, bsx_event_stream* s, size_t* released);
/* Sends the partial chunk and returns once the sink has taken every chunk. */
int bsx_event_stream_finish(bsx_event_stream* s, void* stream);
int bsx_event_stream_stats(const bsx_event_stream* s, bsx_stream_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* BSX_OUTPUT_H_ */
