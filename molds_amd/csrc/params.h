#pragma once
#include "nddo_device.h"

namespace nddo {
// Fills *out with the converted parameters of element Z for theory
// 0 = MNDO, 1 = AM1, 2 = PM3.  Returns -1 for an unsupported element.
int elem_params(int Z, int theory, ElemDev* out);
// sum over atoms of (dHf_atom - E_isol_atom) for the MNDO heat of formation
// (src/mndo/Mndo.cpp:395-404); one atom's contribution.
double mndo_atom_hof_offset(int Z);
// Isolated-atom electronic energy (eV) from the one-centre parameters of a theory.
double eisol_ev(int Z, int theory);
// One atom's dHf_atom - E_isol,atom for the heat of formation of any theory:
// MNDO with MolDS's tabulated E_isol, AM1 / PM3 with eisol_ev (MolDS computes no
// AM1 / PM3 heat of formation, src/am1/Am1.cpp:261-263).
double atom_hof_offset(int Z, int theory);
bool element_supported(int Z);
// Standard atomic weight (g/mol) as MolDS's atom classes set it (Hatom.cpp:52, Catom.cpp:49, ...).
double atomic_mass_gmol(int Z);
}  // namespace nddo
