// Scene and camera records as the path-tracing kernels read them from HBM;
// shared by the host side (api.cpp) and the kernels.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>
#include "layout.h"

namespace illume {

struct DevMaterial {  // illume_material
    int32_t type;
    float kd[3], eta[3], k[3];
    float alpha, ior;
    float pad[4];
};
enum { kMatte = 0, kMetal = 1, kGlass = 2 };

struct DevLight {  // one-sided DiffuseAreaLight on a triangle
    float p0[3], p1[3], p2[3];
    float L[3];
    float area;
    uint32_t prim;
    float pad[2];
};

struct DevScene {
    const GpuNode4Q *nodes;     // four-child quantized BVH
    const PrimRecord *prims;    // leaf order
    const float4 *tri_verts;    // by triangle id: p0, p1, p2 (xyz)
    const float4 *spheres;      // by sphere index: cx, cy, cz, r
    const uint32_t *prim_material;
    const int32_t *prim_light;  // light index or -1
    const DevMaterial *materials;
    const DevLight *lights;
    const float *light_func;    // power distribution (sampling.rs:8-93)
    const float *light_cdf;     // n_lights + 1
    float light_func_int;
    uint32_t n_lights;          // area lights + the sky (if any)
    uint32_t n_tris;
    int32_t sky_index;          // light index of the analytic sky, -1 if none
    float sky_zenith[3], sky_horizon[3];
};

struct DevCamera {
    float raster_to_camera[16];  // row-major 4x4
    float camera_to_world[12];   // row-major 3x4
    uint32_t width, height;
};

}  // namespace illume
