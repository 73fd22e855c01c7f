// Internal host-side declarations of libassist.
#pragma once
#include <stddef.h>

#include "../../include/assist.h"
#include "../../include/spk.h"

namespace assist_gpu { struct DeviceEphemCache; }

namespace assist_host {

// Number of segments of a target (its one/two arrays).
inline int segments(const spk_target& t) { return t.ind + 1; }
int find_target(const spk_s* s, int code);

// Position / velocity / acceleration of an SPK target in km, km/s, km/s^2.
void target_pva(const spk_s* s, const spk_target& t, double jd_ref, double jd_rel, double u[3], double v[3],
                double w[3]);
int asteroid_pos(const spk_s* s, double jd_ref, double jd_rel, int m, double* GM, double out[3]);
int planets_calc(const assist_ephem* e, double jd_ref, double jd_rel, int body, double* GM, double* x, double* y,
                 double* z, double* vx, double* vy, double* vz, double* ax, double* ay, double* az);

}  // namespace assist_host

namespace assist_gpu {
// Device copy of an ephemeris (built on first use, rebuilt when jd_ref changes).
void free_device_ephem(DeviceEphemCache* c);
}
