// Frozen per-pixel sensor maps and the camera's readout, drawn on the host in fp64.
//
// Every map is a fixed function of (seed, camera, pixel) drawn from Philox purpose kSensorMap: block 0
// gives four normals by Box-Muller (fp64 log, sqrt and sincos), block 1 four open uniforms, block 2
// the normal of a hot pixel's log-normal rate (reference: draw_maps, src/sensor/maps.cpp:78-156).
// The values are the reference's, bit for bit, which the unit tests check against the reference's
// own maps.
#pragma once

#include <stdint.h>

#include <vector>

#include "params.h"
#include "readout.h"

namespace bsx {

// Structure of arrays, one entry per pixel (y * width + x).
struct SensorMaps {
  std::vector<double> theta_on, theta_off, leak, dark_fA, phi0;
  std::vector<uint8_t> hot;          // 0: none, 1: leak-type (the only kind without the cascade's noise)
  std::vector<double> hot_rate_hz;   // a hot pixel's target event rate (0: none)
  std::vector<float> gain;    // static photometric gain, 1 unless set_photometry sets a map
  size_t size() const { return theta_on.size(); }
};

SensorMaps draw_sensor_maps(const SensorParams& p, uint64_t seed, uint32_t camera, uint32_t n_pixels,
                            int num_threads = 0);

// The readout of a camera (the counter phase drawn from seed and camera).
Readout readout_of(const SensorParams& p, uint64_t seed, uint32_t camera);

}  // namespace bsx
