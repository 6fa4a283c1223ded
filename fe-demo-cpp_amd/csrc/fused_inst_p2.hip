// Compiled fused-mass shapes with P = 2 (registered at library load).
#include "fused_mass_kernel.hpp"

namespace femhip {
namespace {

const FusedShapeReg<2, 3, 1, 7, 7> reg_2_3_1_7;
const FusedShapeReg<2, 3, 2, 7, 7> reg_2_3_2_7;
const FusedShapeReg<2, 3, 3, 7, 7> reg_2_3_3_7;

}  // namespace
}  // namespace femhip
