// Launch interface of the HIP kernels (internal to libassist).
#pragma once
#include <hip/hip_runtime.h>

#include "device_types.h"

namespace assist_gpu {

// Largest block of the integration kernel.
constexpr int kIas15Block = 256;
// Doubles per LDS column of the integration kernel (b and g, 42 words, padded to 43).
constexpr int kLdsColWords = 43;
// LDS columns per lane group: one per owned vector; 8-lane groups with at most 7 vectors
// (a particle and its 6 variational particles) pack 7 columns.
__host__ __device__ constexpr int ias15_cols_per_group(int G, int nvec) { return (G == 8 && nvec <= 7) ? 7 : G; }

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
// waves: the kernel variant compiled for 1 or 2 waves per SIMD (register budget 512 or 256).
hipError_t launch_ias15(const IntegrateArgs& a, int G, bool mixed, int waves, int blocks, int threads, hipStream_t s);
hipError_t ias15_occupancy(int G, bool mixed, int waves, int nvec, int threads, int* blocks_per_cu);
size_t ias15_lds_bytes(int G, int nvec, int threads);
hipError_t launch_set_counter(int* q, int v, hipStream_t s);

}  // namespace assist_gpu
