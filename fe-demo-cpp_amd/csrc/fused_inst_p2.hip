// Compiled fused-mass shapes with P = 2 (registered at library load).
#include "batch_kernel.hpp"
#include "fused_box_kernel.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<2, 3, 1, 64, 64> reg_2_3_1_64;
const BatchShapeReg<2, 3, 1, 64, 64> batch_2_3_1_64;
const FusedShapeReg<2, 3, 2, 64, 64> reg_2_3_2_64;
const BatchShapeReg<2, 3, 2, 64, 64> batch_2_3_2_64;
const FusedShapeReg<2, 3, 3, 64, 64> reg_2_3_3_64;
const BatchShapeReg<2, 3, 3, 64, 64> batch_2_3_3_64;

const FusedBlockShapeReg<2, 3, 1, 32, 32> blk_2_3_1;
const FusedBlockShapeReg<2, 3, 2, 32, 32> blk_2_3_2;
const FusedBlockShapeReg<2, 3, 3, 32, 32> blk_2_3_3;

const FusedBoxReg<2, 3, 1, 4, 4> box_2_3_1_16;
const FusedBoxReg<2, 3, 2, 4, 2> box_2_3_2_8;
const FusedBoxReg<2, 3, 3, 4, 2> box_2_3_3_8;

// One or two elements per wavefront (fused_box_wave_kernel).
const FusedBoxWaveReg<2, 3, 1, 4, 4, 2> wave2_2_3_1_16;
const FusedBoxWaveReg<2, 3, 2, 4, 2, 2> wave2_2_3_2_8;
const FusedBoxWaveReg<2, 3, 3, 4, 2, 2> wave2_2_3_3_8;

}  // namespace
}  // namespace femhip
