// Compiled fused-mass shapes with P = 5 (registered at library load).
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<5, 6, 1, 7, 7> reg_5_6_1_7;
const FusedShapeReg<5, 6, 1, 16, 16> reg_5_6_1_16;
const FusedShapeReg<5, 6, 2, 7, 7> reg_5_6_2_7;
const FusedShapeReg<5, 6, 3, 7, 7> reg_5_6_3_7;

}  // namespace
}  // namespace femhip
