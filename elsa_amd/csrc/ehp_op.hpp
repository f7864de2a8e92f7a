// ehp_op.hpp -- the library's handles (opaque in include/elsa_hip.h): a geometry, an operator
// and its device state.
#pragma once

#include "ehp_kernels.hpp"
#include "ehp_order.hpp"

#include <vector>

struct ehp_geometry {
    int dim;
    int64_t nposes;
    int vol[3];
    int det[2];
    std::vector<ehp::PoseGeom> host;
    std::vector<float> pix; // curved detectors: planar coordinate of each pixel centre (empty: planar)
};

enum op_kind { KIND_JOSEPH = 0, KIND_SIDDON = 1, KIND_VOXEL = 2 };

struct ehp_op {
    int kind;
    int dtype;
    int device;
    ehp::Dims dims;
    ehp::PoseGeom* d_geo = nullptr;
    void* d_lut = nullptr;
    int n_lut = 0;
    float support = 0.0f;
    int differential = 0; // voxel projectors: phase-contrast weight
    float* d_pix = nullptr; // curved detectors: planar pixel coordinates (dims.pix points here)
    std::vector<float> host_pix;    // the same in host memory
    bool curved_tiles = false;      // curved detectors: the coordinates allow the tiled back projections
    uint64_t samples = 0;
    // Joseph, FP32: plane-major forward and LDS-tiled back projection
    bool planes = false;
    ehp::PadLayout layouts[3]{};
    void* d_view[3] = {nullptr, nullptr, nullptr}; // zero-padded plane-major volume views (op's storage type)
    float* d_acc32 = nullptr;                        // bfloat16 operators: FP32 back projection before rounding
    bool tiled = false;
    int* d_pose_axes = nullptr;
    int axes_present = 0;
    ehp::TileCfg tile{};
    ehp::RayRec* d_table = nullptr;   // per-ray traversal state (Joseph)
    ehp::SidRec* d_sidtable = nullptr; // per-ray traversal state (Siddon)
    float* d_pose_margin = nullptr;    // voxel forward: frustum margin per pose (pixels)
    unsigned int* d_maxbits = nullptr; // per-block maxima of |y| of the current back projection (kAbsmaxBlocks)
    int* d_fixexp = nullptr;           // fixed-point exponent of the current back projection
    double slot_bound = 0.0;
    int batch = ehp::kPoseBatch;
    std::vector<ehp::PoseGeom> host_geo;
    void* d_stage = nullptr; // atomic scatters whose output is not plain device memory go here first
    size_t stage_bytes = 0;
    ehp::OrderGuard order;   // calls on this operator run one after another on the device
};
