/* bsx rigs: several cameras (sensor models of any resolution) advanced by one launch chain.
 *
 * The reference engine advances its cameras one after the other on one stream, one
 * SensorModel::advance_device call each (src/engine/engine.h:1-8, src/engine/engine_impl.h:1343-1348).
 * A rig replaces that per-camera loop: one call runs phase A over every camera's segments, one
 * heavy-first order of the joint pixel space and one walk launch over all pixels (persistent, fed by
 * a work queue, when the tuning's queue chunk is positive), captured once into a hipGraph and replayed
 * while the calls' extents fit what was captured. Every camera's events, state and statistics are
 * those of advancing it alone with bsx_sensor_advance_device.
 *
 * Conventions are those of bsx_sensor.h (status codes, bsx_last_error, int64 ns, the caller's
 * stream). The rig borrows its cameras: each must outlive it and must not be advanced elsewhere
 * while a rig call runs.
 */
#ifndef BSX_RIG_H_
#define BSX_RIG_H_

#include "bsx_sensor.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bsx_rig bsx_rig;

typedef struct bsx_rig_stats_t {
  uint64_t launches;       /* rig calls launched */
  uint64_t graph_launches; /* of which as one graph replay */
  uint64_t captures;       /* graphs captured (first call, new tuning, calls that outgrew the capacities) */
  uint64_t walk_reruns;    /* walks re-run after an event-buffer overflow */
  double kernel_ms;        /* summed device time of the rig calls */
} bsx_rig_stats_t;

/* No reference counterpart (the engine's camera loop, above). 1 to 64 distinct cameras sharing one
 * photoreceptor kind (BSX_E_INVALID otherwise). */
int bsx_rig_create(bsx_sensor* const* cameras, int32_t n_cameras, bsx_rig** out);
int bsx_rig_destroy(bsx_rig* rig);

/* Advances camera c from its own model time to t_end_ns on runs[c] (the reference's DeviceRuns
 * layout, src/sensor/sensor_model.h:81-89; n = 0 holds the light), and returns its unordered device
 * events in out[c] (valid until that camera's next advance). Every camera's input checks pass
 * before any camera changes (BSX_E_INVALID names the camera); a failure after them leaves every
 * camera uninitialised. Reference per camera: advance_device(runs, t_end, stream),
 * src/sensor/sensor_model.h:166. */
int bsx_rig_advance_device(bsx_rig* rig, const bsx_device_runs* runs, int64_t t_end_ns, void* stream,
                           bsx_device_events* out);

/* The walk's launch shape (the fields of bsx_launch_policy other than event_capacity; 0 in a field =
 * keep, queue_chunk -1 = no queue). A new shape re-captures the graph at the next call. */
int bsx_rig_set_tuning(bsx_rig* rig, const bsx_launch_policy* tuning);
int bsx_rig_get_tuning(bsx_rig* rig, bsx_launch_policy* out);
/* Graph capture on (default) or off (kernel-by-kernel launches; same results). */
int bsx_rig_set_graph(bsx_rig* rig, int32_t on);
int bsx_rig_stats(bsx_rig* rig, bsx_rig_stats_t* out);

/* The autotuner: runs phase A and the order of the given call once, then times the walk alone under
 * candidate shapes (coordinate descent over block size {64, 128, 256}, walk budget {4 .. 64},
 * search batch {1 .. 64} and queue chunk {0 .. 16}, two passes from the current shape; `reps` timed
 * walks per candidate, minimum kept) and keeps the fastest as the rig's tuning. No camera's state or
 * time changes (the cameras' event buffers are overwritten). Up to max_trials trials are returned
 * (tunings[i], walk_ms[i]; either pointer may be null), their count in *n_trials. */
int bsx_rig_autotune(bsx_rig* rig, const bsx_device_runs* runs, int64_t t_end_ns, void* stream, int32_t reps,
                     bsx_launch_policy* tunings, double* walk_ms, int32_t max_trials, int32_t* n_trials);

#ifdef __cplusplus
}
#endif

#endif /* BSX_RIG_H_ */
