// CellGradient worklet (dax/worklet/CellGradient.h:28-48):
// ControlSignature(TopologyIn, FieldPointIn, FieldPointIn, FieldOut),
// ExecutionSignature _4(_2,_3).  Executed by dax_hip_cell_gradient.
#ifndef DAX_HIP_SLICE_WORKLET_CELLGRADIENT_H
#define DAX_HIP_SLICE_WORKLET_CELLGRADIENT_H

namespace dax {
namespace worklet {
class CellGradient {};
}  // namespace worklet
}  // namespace dax

#endif
