// Compiled fused-mass shapes with P = 6 (registered at library load).
#include "batch_kernel.hpp"
#include "fused_box_kernel.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<6, 7, 1, 9, 9> reg_6_7_1_9;
const BatchShapeReg<6, 7, 1, 9, 9> batch_6_7_1_9;
const FusedShapeReg<6, 7, 1, 5, 5> reg_6_7_1_5;
const BatchShapeReg<6, 7, 1, 5, 5> batch_6_7_1_5;
const FusedShapeReg<6, 7, 2, 9, 9> reg_6_7_2_9;
const BatchShapeReg<6, 7, 2, 9, 9> batch_6_7_2_9;
const FusedShapeReg<6, 7, 3, 9, 9> reg_6_7_3_9;
const BatchShapeReg<6, 7, 3, 9, 9> batch_6_7_3_9;

const FusedBlockShapeReg<6, 7, 1, 8, 8> blk_6_7_1;
const FusedBlockShapeReg<6, 7, 2, 4, 4> blk_6_7_2;
const FusedBlockShapeReg<6, 7, 3, 4, 4> blk_6_7_3;

const FusedBoxReg<6, 7, 1, 4, 2> box_6_7_1_8;
const FusedBoxReg<6, 7, 1, 2, 2> box_6_7_1_4;
const FusedBoxReg<6, 7, 2, 2, 2> box_6_7_2_4;
const FusedBoxReg<6, 7, 3, 2, 2> box_6_7_3_4;

// One or two elements per wavefront (fused_box_wave_kernel).
const FusedBoxWaveReg<6, 7, 1, 4, 2> wave_6_7_1_8;
const FusedBoxWaveReg<6, 7, 1, 2, 2> wave_6_7_1_4;
const FusedBoxWaveReg<6, 7, 2, 2, 2> wave_6_7_2_4;
const FusedBoxWaveReg<6, 7, 3, 2, 2> wave_6_7_3_4;

}  // namespace
}  // namespace femhip
