// Evaluation modes of the element kernels (the `mode` of
// tad_device_element_args).
#pragma once

namespace tad {

enum EvalMode
{
    MODE_GRAD = 1,      // value + gradient
    MODE_HESS = 2,      // value + gradient + Hessian
    MODE_HESS_PROJ = 3, // value + gradient + projected Hessian
};

} // namespace tad
