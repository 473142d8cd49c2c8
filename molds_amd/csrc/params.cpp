// Element parameters for MNDO / AM1 / PM3 (H, C, N, O, F).
//
// Values are the published parameter sets as MolDS lists them
// (src/base/atoms/Hatom.cpp:50-220, Catom.cpp:47-225, Natom.cpp, Oatom.cpp:47-225,
// Fatom.cpp; literature keys [DT_1977] MNDO, [DZHS_1985] AM1, [S_1989] PM3), in
// eV / Angstrom units, converted to atomic units with MolDS's constants
// (src/base/Parameters.cpp:45-48) the way the reference accessors do
// (src/base/atoms/Atom.cpp:263-1109).
#include <cmath>
#include <cstring>

#include "nddo_device.h"
#include "params.h"

namespace nddo {

namespace {

const double kEV2AU = 0.03674903;
const double kKcal2AU = 0.00159360175;
const double kAng2AU = 1.0 / 0.5291772;

struct Published {
    // U_ss U_pp zeta_s zeta_p beta_s beta_p alpha(1/A)
    double uss, upp, zs, zp, bs, bp, alpha;
    // one-centre integrals (eV)
    double gss, gpp, gsp, gpp2, hsp;
    // D1 D2 (bohr) and the inverse additive terms 0.5/rho_k (bohr^-1)
    double d1, d2, inv_rho0, inv_rho1, inv_rho2;
    double k[4], l[4], m[4];   // eV, A^-2, A
};

struct Element {
    int Z, nao, nshell;
    double core;
    double mndo_eisol_ev, mndo_dhf_kcal;   // MNDO atomic energy, heat of formation
    Published mndo, am1, pm3;
};

const Element kElements[] = {
    {1, 1, 1, 1.0, -11.906276, 52.102,
     {-11.906276, 0, 1.331967, 0, -6.989064, 0, 2.544134, 12.848, 0, 0, 0, 0,
      0, 0, 0.4721515374, 0, 0, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
     {-11.396427, 0, 1.188078, 0, -6.173787, 0, 2.882324, 12.848, 0, 0, 0, 0,
      0, 0, 0.4721515374, 0, 0, {0.122796, 0.005090, -0.018336, 0}, {5.0, 5.0, 2.0, 0}, {1.2, 1.8, 2.1, 0}},
     {-13.073321, 0, 0.967807, 0, -5.626512, 0, 3.356386, 14.794208, 0, 0, 0, 0,
      0, 0, 0.5436727936, 0, 0, {1.128750, -1.060329, 0, 0}, {5.096282, 6.003788, 0, 0}, {1.537465, 1.570189, 0, 0}}},
    {6, 4, 2, 4.0, -120.500606, 170.89,
     {-52.279745, -39.205558, 1.787537, 1.787537, -18.985044, -7.934122, 2.546380, 12.23, 11.08, 11.47, 9.84, 2.43,
      0.8074661800, 0.6851577737, 0.4494406369, 0.6149309919, 0.6685771472, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
     {-52.028658, -39.614239, 1.808665, 1.685116, -15.715783, -7.719283, 2.648274, 12.23, 11.08, 11.47, 9.84, 2.43,
      0.8236735591, 0.7268015207, 0.4494406369, 0.6082783276, 0.6423370115,
      {0.011355, 0.045924, -0.020061, -0.001260}, {5.0, 5.0, 5.0, 5.0}, {1.60, 1.85, 2.05, 2.65}},
     {-47.270320, -36.266918, 1.565085, 1.842345, -11.910015, -9.802755, 2.707807,
      11.200708, 10.796292, 10.265027, 9.042566, 2.290980,
      0.8332396384, 0.6647749859, 0.4116151543, 0.5885706542, 0.7647513703,
      {0.050107, 0.050733, 0, 0}, {6.003165, 6.002979, 0, 0}, {1.642214, 0.892488, 0, 0}}},
    {7, 4, 2, 5.0, -202.581201, 113.00,
     {-71.932122, -57.172319, 2.255614, 2.255614, -20.495758, -20.495758, 2.861342, 13.59, 12.98, 12.66, 11.59, 3.14,
      0.6399036683, 0.5429762678, 0.4994193177, 0.7843433156, 0.8126295047, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
     {-71.860000, -57.167581, 2.315410, 2.157940, -20.299110, -18.238666, 2.947286, 13.59, 12.98, 12.66, 11.59, 3.14,
      0.6433247425, 0.5675527917, 0.4994193177, 0.7820630445, 0.7883351388,
      {0.025251, 0.028953, -0.005806, 0}, {5.0, 5.0, 2.0, 0}, {1.50, 2.10, 2.40, 0}},
     {-49.335672, -47.509736, 2.028094, 2.313728, -14.062521, -20.043848, 2.830545,
      11.904787, 11.754672, 7.348565, 10.807277, 1.136713,
      0.6577005762, 0.5293383109, 0.4374893746, 0.5030877737, 0.7364801616,
      {1.501674, -1.505772, 0, 0}, {5.901148, 6.004658, 0, 0}, {1.710740, 1.716149, 0, 0}}},
    {8, 4, 2, 6.0, -317.868506, 59.559,
     {-99.64309, -77.797472, 2.699905, 2.699905, -32.688082, -32.688082, 3.160604, 15.42, 14.52, 14.48, 12.98, 3.94,
      0.5346023927, 0.4536251725, 0.5666700426, 0.9592303457, 0.9495760934, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
     {-97.830000, -78.262380, 3.108032, 2.524039, -29.272773, -29.272773, 4.455371, 15.42, 14.52, 14.48, 12.98, 3.94,
      0.4988896404, 0.4852321503, 0.5666700426, 0.9960801167, 0.9065055775,
      {0.280962, 0.081430, 0, 0}, {5.0, 7.0, 0, 0}, {0.847918, 1.445071, 0, 0}},
     {-86.993002, -71.879580, 3.796544, 2.389402, -45.202651, -24.752515, 3.217102,
      15.755760, 13.654016, 10.621160, 12.40609, 0.593883,
      0.4086173087, 0.5125738036, 0.5790088969, 0.5299517372, 0.8179482975,
      {-1.131128, 1.137891, 0, 0}, {6.002477, 5.950512, 0, 0}, {1.607311, 1.598395, 0, 0}}},
    {9, 4, 2, 7.0, -476.683781, 18.86,
     {-131.071548, -105.782137, 2.848487, 2.848487, -48.290460, -36.508540, 3.419661, 16.92, 16.71, 17.25, 14.91, 4.83,
      0.5067166088, 0.4299633003, 0.6217935876, 1.0850000098, 1.0343451703, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
     {-136.105579, -104.889885, 3.770082, 2.494670, -69.590277, -27.922360, 5.517800, 16.92, 16.71, 17.25, 14.91, 4.83,
      0.4145203025, 0.4909446425, 0.6217935876, 1.2088469198, 0.9449175360,
      {0.242079, 0.003607, 0, 0}, {4.80, 4.60, 0, 0}, {0.930, 1.660, 0, 0}},
     {-110.435303, -105.685047, 4.708555, 2.491178, -48.405939, -27.744660, 3.358921,
      10.496667, 14.817256, 16.073689, 14.418393, 0.727763,
      0.3125302275, 0.4916328225, 0.3857423305, 0.6768359077, 0.6119953427,
      {-0.012166, -0.002852, 0, 0}, {6.023574, 6.003717, 0, 0}, {1.856859, 2.636158, 0, 0}}},
};

const Element* find(int Z) {
    for (const Element& e : kElements)
        if (e.Z == Z) return &e;
    return nullptr;
}

}  // namespace

int elem_params(int Z, int theory, ElemDev* out) {
    const Element* e = find(Z);
    if (!e || theory < 0 || theory > 2) return -1;
    const Published& p = theory == 0 ? e->mndo : (theory == 1 ? e->am1 : e->pm3);
    std::memset(out, 0, sizeof(*out));
    out->Z = Z;
    out->nao = e->nao;
    out->nshell = e->nshell;
    out->core = e->core;
    out->U[0] = p.uss * kEV2AU;
    out->U[1] = p.upp * kEV2AU;
    out->zeta[0] = p.zs;
    out->zeta[1] = p.zp;
    out->beta[0] = p.bs * kEV2AU;
    out->beta[1] = p.bp * kEV2AU;
    out->alpha = p.alpha / kAng2AU;
    out->Gss = p.gss * kEV2AU;
    out->Gpp = p.gpp * kEV2AU;
    out->Gsp = p.gsp * kEV2AU;
    out->Gpp2 = p.gpp2 * kEV2AU;
    out->Hsp = p.hsp * kEV2AU;
    out->Hpp = 0.5 * (out->Gpp - out->Gpp2);   // [MOPAC_1990] p17
    out->D[0] = 0.0;
    out->D[1] = p.d1;
    out->D[2] = p.d2;
    out->rho[0] = 0.5 / p.inv_rho0;
    out->rho[1] = p.inv_rho1 != 0.0 ? 0.5 / p.inv_rho1 : 0.0;
    out->rho[2] = p.inv_rho2 != 0.0 ? 0.5 / p.inv_rho2 : 0.0;
    for (int k = 0; k < 4; ++k) {
        out->K[k] = p.k[k] * kEV2AU;
        out->L[k] = p.l[k] / std::pow(kAng2AU, 2.0);
        out->M[k] = p.m[k] * kAng2AU;
    }
    return 0;
}

double mndo_atom_hof_offset(int Z) {
    const Element* e = find(Z);
    if (!e) return 0.0;
    return -e->mndo_eisol_ev * kEV2AU + e->mndo_dhf_kcal * kKcal2AU;
}

// Electronic energy of the isolated atom in its ground configuration s^ns p^np
// (Hund's rule), from the one-centre parameters (Dewar & Thiel, J. Am. Chem.
// Soc. 99, 4899 (1977), the EISOL of MOPAC):
//   E = ns U_ss + np U_pp + [ns = 2] G_ss + ns np G_sp - (ns np / 2) H_sp
//       + a G_pp + b G_p2,
// with (a, b) = (0, 0), (0, 0), (-1/2, 3/2), (-3/2, 9/2), (-1/2, 13/2), (0, 10)
// for np = 0..5 (the p-shell pairs' Coulomb G_p2 and exchange
// H_pp = (G_pp - G_p2) / 2 counted for the high-spin configuration).  It gives
// the published AM1 and PM3 atomic energies (e.g. AM1 C -120.815794 eV,
// PM3 O -289.342 eV) and MNDO's for H, C and F.
double eisol_ev(int Z, int theory) {
    const Element* e = find(Z);
    if (!e || theory < 0 || theory > 2) return 0.0;
    const Published& p = theory == 0 ? e->mndo : (theory == 1 ? e->am1 : e->pm3);
    const int ns = Z == 1 ? 1 : 2, np = Z == 1 ? 0 : (int)e->core - 2;
    static const double a[6] = {0.0, 0.0, -0.5, -1.5, -0.5, 0.0}, b[6] = {0.0, 0.0, 1.5, 4.5, 6.5, 10.0};
    return ns * p.uss + np * p.upp + (ns == 2 ? p.gss : 0.0) + ns * np * p.gsp - 0.5 * ns * np * p.hsp +
           a[np] * p.gpp + b[np] * p.gpp2;
}

double atom_hof_offset(int Z, int theory) {
    const Element* e = find(Z);
    if (!e) return 0.0;
    if (theory == 0) return mndo_atom_hof_offset(Z);   // MolDS's MNDO tables (Mndo.cpp:395-404)
    return -eisol_ev(Z, theory) * kEV2AU + e->mndo_dhf_kcal * kKcal2AU;
}

bool element_supported(int Z) { return find(Z) != nullptr; }

double atomic_mass_gmol(int Z) {
    switch (Z) {
        case 1: return 1.00794;
        case 6: return 12.0107;
        case 7: return 14.00674;
        case 8: return 15.9994;
        case 9: return 18.9984032;
        default: return 0.0;
    }
}

}  // namespace nddo
