// Launch interface of the HIP kernels (internal to libassist).
#pragma once
#include <hip/hip_runtime.h>

#include "device_types.h"

namespace assist_gpu {

struct ForcesArgs {
    const DevEphem* E;
    ForceOpts fo;
    int nsys;
    int nvec;               // 1 + number of variational particles per system
    const double* t;        // [nsys]
    const double* states;   // [nsys][nvec][6]
    const double* params;   // [nsys][nvec][3] or nullptr
    double* acc;            // [nsys][nvec][3]
    int* status;            // [nsys] or nullptr
};

struct IntegrateArgs {
    const DevEphem* E;
    ForceOpts fo;
    IntegratorOpts io;
    int nsys;
    int nvec;
    const double* t0;       // [nsys]
    const double* t1;       // [nsys]
    double* states;         // [nsys][nvec][6], updated in place
    const double* params;   // [nsys][nvec][3] or nullptr
    int* status;            // [nsys]
    long long* steps;       // [nsys] or nullptr
    long long* evals;       // [nsys] or nullptr
    long long* rejected;    // [nsys] or nullptr
    double* t_out;          // [nsys] or nullptr
    double* scratch;        // [threads][42]: br, er of each lane's last accepted step
    int* queue;             // next system to hand out
};

hipError_t launch_forces(const ForcesArgs& a, int G, bool mixed, hipStream_t s);
hipError_t launch_ias15(const IntegrateArgs& a, int G, bool mixed, int blocks, int threads, hipStream_t s);
hipError_t ias15_occupancy(int G, bool mixed, int threads, int* blocks_per_cu);
hipError_t launch_set_counter(int* q, int v, hipStream_t s);

}  // namespace assist_gpu
