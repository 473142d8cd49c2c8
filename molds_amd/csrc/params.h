#pragma once
#include "nddo_device.h"

namespace nddo {
// Fills *out with the converted parameters of element Z for theory
// 0 = MNDO, 1 = AM1, 2 = PM3.  Returns -1 for an unsupported element.
int elem_params(int Z, int theory, ElemDev* out);
// sum over atoms of (dHf_atom - E_isol_atom) for the MNDO heat of formation
// (src/mndo/Mndo.cpp:395-404); one atom's contribution.
double mndo_atom_hof_offset(int Z);
bool element_supported(int Z);
}  // namespace nddo
