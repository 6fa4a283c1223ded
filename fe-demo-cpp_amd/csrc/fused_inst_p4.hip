// Compiled fused-mass shapes with P = 4 (registered at library load).
#include "batch_kernel.hpp"
#include "fused_box_kernel.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<4, 5, 1, 23, 23> reg_4_5_1_23;
const BatchShapeReg<4, 5, 1, 23, 23> batch_4_5_1_23;
const FusedShapeReg<4, 5, 1, 10, 10> reg_4_5_1_10;
const BatchShapeReg<4, 5, 1, 10, 10> batch_4_5_1_10;
const FusedShapeReg<4, 5, 2, 23, 23> reg_4_5_2_23;
const BatchShapeReg<4, 5, 2, 23, 23> batch_4_5_2_23;
const FusedShapeReg<4, 5, 3, 23, 23> reg_4_5_3_23;
const BatchShapeReg<4, 5, 3, 23, 23> batch_4_5_3_23;

const FusedBlockShapeReg<4, 5, 1, 16, 16> blk_4_5_1;
const FusedBlockShapeReg<4, 5, 2, 8, 8> blk_4_5_2;
const FusedBlockShapeReg<4, 5, 3, 8, 8> blk_4_5_3;

const FusedBoxReg<4, 5, 1, 4, 4> box_4_5_1_16;
const FusedBoxReg<4, 5, 1, 2, 2> box_4_5_1_4;
const FusedBoxReg<4, 5, 2, 2, 2> box_4_5_2_4;
const FusedBoxReg<4, 5, 3, 2, 2> box_4_5_3_4;

// One or two elements per wavefront (fused_box_wave_kernel).
const FusedBoxWaveReg<4, 5, 1, 4, 4, 2> wave2_4_5_1_16;
const FusedBoxWaveReg<4, 5, 2, 2, 2, 2> wave2_4_5_2_4;
const FusedBoxWaveReg<4, 5, 3, 2, 2, 2> wave2_4_5_3_4;

}  // namespace
}  // namespace femhip
