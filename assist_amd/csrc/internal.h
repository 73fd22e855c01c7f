// Internal host-side declarations of libassist.
#pragma once
#include <stddef.h>

#include <string>
#include <vector>

#include "../../include/assist.h"

struct spk_target {
    int code;
    int cen;
    double mass;  // GM; 0 if no constant was found
    double beg;   // JD of the first segment's start
    double end;   // JD of the last segment's end
    double res;   // length of the first segment in days
    std::vector<int> one, two;  // 1-based word addresses of each segment
};

namespace assist_gpu { struct DeviceEphemCache; }

struct spk_s {
    std::vector<spk_target> targets;
    int num = 0;
    void* map = nullptr;
    size_t len = 0;
    assist_gpu::DeviceEphemCache* gpu = nullptr;  // device copy (planets file only)
};

struct spk_constants {
    double AU = 0, EMRAT = 0, J2E = 0, J3E = 0, J4E = 0, J2SUN = 0, RE = 0, CLIGHT = 0, ASUN = 0;
    std::vector<std::string> names;
    std::vector<double> values;
};

namespace assist_host {

spk_s* spk_init(const char* path);
void spk_free(spk_s* s);
spk_constants load_constants(const char* path);
void join_masses(spk_s* s, const spk_constants& c, double emrat);
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
