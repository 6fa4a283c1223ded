// Compiled fused-mass shapes with P = 8 (registered at library load).
#include "fused_box_kernel.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<8, 10, 1, 3, 3> reg_8_10_1_3;
const FusedShapeReg<8, 10, 1, 2, 2> reg_8_10_1_2;
const FusedShapeReg<8, 10, 3, 3, 3> reg_8_10_3_3;
const FusedShapeReg<8, 10, 3, 2, 2> reg_8_10_3_2;

const FusedBlockShapeReg<8, 10, 1, 3, 3> blk_8_10_1;
const FusedBlockShapeReg<8, 10, 3, 2, 2> blk_8_10_3;

const FusedBoxReg<8, 10, 1, 2, 2> box_8_10_1_4;
const FusedBoxReg<8, 10, 3, 2, 2> box_8_10_3_4;
const FusedBoxSplitReg<8, 10, 3, 4, 2> box_8_10_3_8_split;

// One or two elements per wavefront (fused_box_wave_kernel).
const FusedBoxWaveReg<8, 10, 1, 2, 2> wave_8_10_1_4;
const FusedBoxWaveReg<8, 10, 3, 4, 2> wave_8_10_3_8;
const FusedBoxWaveReg<8, 10, 3, 8, 2> wave_8_10_3_18;
const FusedBoxWaveReg<8, 10, 3, 2, 2> wave_8_10_3_4;

}  // namespace
}  // namespace femhip
