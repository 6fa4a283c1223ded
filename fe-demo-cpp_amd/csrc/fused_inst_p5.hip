// Compiled fused-mass shapes with P = 5 (registered at library load).
#include "batch_kernel.hpp"
#include "fused_box_kernel.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<5, 6, 1, 16, 16> reg_5_6_1_16;
const BatchShapeReg<5, 6, 1, 16, 16> batch_5_6_1_16;
const FusedShapeReg<5, 6, 1, 7, 7> reg_5_6_1_7;
const BatchShapeReg<5, 6, 1, 7, 7> batch_5_6_1_7;
const FusedShapeReg<5, 6, 1, 14, 14> reg_5_6_1_14;
const BatchShapeReg<5, 6, 1, 14, 14> batch_5_6_1_14;
const FusedShapeReg<5, 6, 2, 16, 16> reg_5_6_2_16;
const BatchShapeReg<5, 6, 2, 16, 16> batch_5_6_2_16;
const FusedShapeReg<5, 6, 3, 16, 16> reg_5_6_3_16;
const BatchShapeReg<5, 6, 3, 16, 16> batch_5_6_3_16;

const FusedBlockShapeReg<5, 6, 1, 16, 16> blk_5_6_1;
const FusedBlockShapeReg<5, 6, 1, 14, 14> blk_5_6_1_14;
const FusedBlockShapeReg<5, 6, 1, 8, 8> blk_5_6_1_8;
const FusedBlockShapeReg<5, 6, 1, 7, 7> blk_5_6_1_7;
const FusedBlockShapeReg<5, 6, 2, 8, 8> blk_5_6_2;
const FusedBlockShapeReg<5, 6, 3, 8, 8> blk_5_6_3;

const FusedBoxReg<5, 6, 1, 4, 4> box_5_6_1_16;
const FusedBoxReg<5, 6, 1, 4, 2> box_5_6_1_8;
const FusedBoxReg<5, 6, 1, 2, 2> box_5_6_1_4;
const FusedBoxReg<5, 6, 2, 4, 2> box_5_6_2_8;
const FusedBoxReg<5, 6, 3, 4, 2> box_5_6_3_8;

// One or two elements per wavefront (fused_box_wave_kernel).
const FusedBoxWaveReg<5, 6, 1, 4, 4, 2> wave2_5_6_1_16;
const FusedBoxWaveReg<5, 6, 1, 4, 2, 2> wave2_5_6_1_8;
const FusedBoxWaveReg<5, 6, 1, 8, 2, 2> wave2_5_6_1_18;
const FusedBoxWaveReg<5, 6, 1, 8, 4, 2> wave2_5_6_1_32;
const FusedBoxWaveReg<5, 6, 1, 16, 2, 2> wave2_5_6_1_34;
const FusedBoxWaveReg<5, 6, 2, 4, 2, 2> wave2_5_6_2_8;
const FusedBoxWaveReg<5, 6, 3, 4, 2, 2> wave2_5_6_3_8;

}  // namespace
}  // namespace femhip
