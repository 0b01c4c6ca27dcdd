// mlp.h -- device buffers and launch wrappers of the learned warm start.
#pragma once

#include "dfsph_internal.h"

namespace dfsph {

// Network 16 -> 64 -> 64 -> 64 -> 2; flat FP32 parameter vector
// [W1 (16x64, row = input), b1, W2 (64x64), b2, W3, b3, W4 (64x2), b4].
constexpr int kNF = 16, kNH = 64, kNO = 2;
constexpr int kOW1 = 0, kOB1 = 1024, kOW2 = 1088, kOB2 = 5184, kOW3 = 5248, kOB3 = 9344,
              kOW4 = 9408, kOB4 = 9536, kNParam = 9538;

struct MlpDev {
    float* theta = nullptr;    // FP32 master parameters
    void* wpack = nullptr;     // BF16 MFMA operands packed from theta
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* ds_x = nullptr;     // dataset features [rows][16]
    float* ds_t = nullptr;     // dataset targets  [rows][2]
    uint32_t cus = 0;          // compute units of the device (inference grid)
};

size_t mlp_pack_bytes();
int mlp_smem_ok();
int mlp_set_smem_limits();
void launch_mlp_warm(const SimConst& k, const Buffers& b, const MlpDev& m, int phase, void* stream);
void launch_record_features(const SimConst& k, const Buffers& b, const MlpDev& m, int phase,
                            uint32_t step, uint32_t stride, uint32_t base, uint32_t cap, void* stream);
void launch_record_targets(const SimConst& k, const Buffers& b, const MlpDev& m, uint32_t step,
                           uint32_t stride, uint32_t base, uint32_t cap, void* stream);
void launch_mlp_pack(const MlpDev& m, void* stream);
// tau_density: expectile of the density output's loss (0.5 = MSE).
void launch_mlp_train(const MlpDev& m, const uint32_t* rows, uint32_t nrows, float* part,
                      float* part_loss, float tau_density, void* stream);
void launch_mlp_adam(const MlpDev& m, const float* part, uint32_t nparts, float lr, float b1, float b2,
                     float eps, float c1, float c2, float* grad_out, void* stream);
void launch_mlp_eval(const MlpDev& m, const float* x, uint32_t n, float* out, void* stream);

}  // namespace dfsph
