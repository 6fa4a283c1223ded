// Compiled fused-mass shapes with P = 4 (registered at library load).
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<4, 5, 1, 10, 10> reg_4_5_1_10;
const FusedShapeReg<4, 5, 2, 10, 10> reg_4_5_2_10;
const FusedShapeReg<4, 5, 3, 10, 10> reg_4_5_3_10;

}  // namespace
}  // namespace femhip
