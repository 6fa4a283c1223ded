// Compiled fused-mass shapes with P = 3 (registered at library load).
#include "batch_kernel.hpp"
#include "fused_box_kernel.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<3, 4, 1, 16, 16> reg_3_4_1_16;
const BatchShapeReg<3, 4, 1, 16, 16> batch_3_4_1_16;
const FusedShapeReg<3, 4, 1, 32, 32> reg_3_4_1_32;
const BatchShapeReg<3, 4, 1, 32, 32> batch_3_4_1_32;
const FusedShapeReg<3, 4, 2, 16, 16> reg_3_4_2_16;
const BatchShapeReg<3, 4, 2, 16, 16> batch_3_4_2_16;
const FusedShapeReg<3, 4, 3, 16, 16> reg_3_4_3_16;
const BatchShapeReg<3, 4, 3, 16, 16> batch_3_4_3_16;

const FusedBlockShapeReg<3, 4, 1, 16, 16> blk_3_4_1;
const FusedBlockShapeReg<3, 4, 2, 16, 16> blk_3_4_2;
const FusedBlockShapeReg<3, 4, 3, 16, 16> blk_3_4_3;

const FusedBoxReg<3, 4, 1, 4, 4> box_3_4_1_16;
const FusedBoxReg<3, 4, 1, 2, 2> box_3_4_1_4;
const FusedBoxReg<3, 4, 2, 4, 2> box_3_4_2_8;
const FusedBoxReg<3, 4, 3, 4, 2> box_3_4_3_8;

// One or two elements per wavefront (fused_box_wave_kernel).
const FusedBoxWaveReg<3, 4, 1, 4, 4, 2> wave2_3_4_1_16;
const FusedBoxWaveReg<3, 4, 2, 4, 2, 2> wave2_3_4_2_8;
const FusedBoxWaveReg<3, 4, 3, 4, 2, 2> wave2_3_4_3_8;

}  // namespace
}  // namespace femhip
