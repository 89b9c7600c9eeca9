// Sensor parameters of one camera, the presets of the three sensor families, and validation.
//
// The fields, their meaning, units and presets are the reference's SensorParams
// (src/sensor/params.h:65-172, families src/sensor/param_fields.cpp:363-515): a caller's parameter
// set means the same thing here. validate() applies the reference's cross-key rules
// (src/sensor/params.cpp:44-96) and the rows' hard ranges, and additionally refuses the features
// this backend does not run (it names them), so that an unsupported configuration fails before any
// device work rather than running something else.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../../include/bsx/bsx_sensor.h"

namespace bsx {

using SensorParams = bsx_params;

// The three families of the reference: "prophesee_gen3", "samsung_gen3", "ideal".
SensorParams preset(const std::string& name);
std::vector<std::string> preset_names();

// Throws std::invalid_argument naming the offending field: the reference's rules (validate_ranges),
// then the features this backend does not run (validate).
void validate_ranges(const SensorParams& p);
void validate(const SensorParams& p);

inline bool noise_on(const SensorParams& p) { return p.noise != 0 && p.photoreceptor == BSX_PHOTORECEPTOR_CASCADE; }

// Arrhenius factor exp(-Ea / kB (1/T - 1/T0)) in fp64 (the maps' temperature scaling).
double arrhenius(double ea_eV, double temperature_K, double reference_K);
double median_dark_fA(const SensorParams& p);

}  // namespace bsx
