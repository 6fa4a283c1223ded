// Compiled fused-mass shapes with P = 6 (registered at library load).
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<6, 7, 1, 5, 5> reg_6_7_1_5;
const FusedShapeReg<6, 7, 2, 5, 5> reg_6_7_2_5;
const FusedShapeReg<6, 7, 3, 5, 5> reg_6_7_3_5;

}  // namespace
}  // namespace femhip
