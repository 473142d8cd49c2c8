// Drives the C++ ElectronicStructure-shaped adapter (include/nddo_es.hpp) the
// way MolDS drives Cndo2: SetMolecule, DoSCF, the getters, DoSCF(false) from
// the previous density, and the reference's exceptions.
//
// usage: test_es_adapter <case file>
//   line 1: theory(0 MNDO, 1 AM1, 2 PM3) max_iter rms damping_thresh damping_weight diis_k diis_start diis_end
//   line 2: natoms
//   then natoms lines "Z x y z" (bohr)
// Prints "energy <E> core <Ec> iterations <it> nao <n> eps0 <e0> epsN <eN>" and
// "warm_iterations <k> warm_energy <E>", "forces <3 N values>" (GetForce(0),
// Hartree/bohr), exits 0 when the exception checks pass.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <string>

#include "nddo_es.hpp"

static int expect_throw(const char* what, const std::string& needle, void (*fn)(nddo::NddoElectronicStructure&, nddo::Molecule&),
                        nddo::NddoElectronicStructure& es, nddo::Molecule& m) {
    try {
        fn(es, m);
    } catch (const nddo::MolDSException& ex) {
        if (std::string(ex.what()).find(needle) != std::string::npos) return 0;
        std::printf("FAIL %s: unexpected message: %s\n", what, ex.what());
        return 1;
    }
    std::printf("FAIL %s: no exception\n", what);
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::ifstream in(argv[1]);
    int theory, max_iter, diis_k, natoms;
    double rms, dth, dw, ds, de;
    in >> theory >> max_iter >> rms >> dth >> dw >> diis_k >> ds >> de >> natoms;
    nddo::Molecule mol;
    if (!in || natoms <= 0) {
        std::printf("bad case file header\n");
        return 2;
    }
    for (int a = 0; a < natoms; ++a) {
        int z = 0;
        double x = 0.0, y = 0.0, zz = 0.0;
        if (!(in >> z >> x >> y >> zz)) {
            std::printf("bad case file: atom %d\n", a);
            return 2;
        }
        mol.AddAtom(z, x, y, zz);
    }
    auto engine = std::make_shared<nddo::Engine>(0);
    nddo::NddoElectronicStructure es((nddo::TheoryType)theory, engine);
    es.Parameters().SetMaxIterationsSCF(max_iter);
    es.Parameters().SetThresholdSCF(rms);
    es.Parameters().SetDampingThreshSCF(dth);
    es.Parameters().SetDampingWeightSCF(dw);
    es.Parameters().SetDiisNumErrorVectSCF(diis_k);
    es.Parameters().SetDiisStartErrorSCF(ds);
    es.Parameters().SetDiisEndErrorSCF(de);
    es.SetMolecule(&mol);
    es.DoSCF();
    const int n = es.GetTotalNumberAOs();
    double const* const* C = es.GetFockMatrix();
    double const* eps = es.GetEnergiesMO();
    int fails = 0;
    for (int i = 0; i < n; ++i) {   // rows are MOs with a non-negative element sum (Lapack.cpp:139-150)
        double s = 0.0, nrm = 0.0;
        for (int j = 0; j < n; ++j) { s += C[i][j]; nrm += C[i][j] * C[i][j]; }
        if (s < -1e-12 || std::fabs(nrm - 1.0) > 1e-10) { std::printf("FAIL row %d sum %g norm %g\n", i, s, nrm); ++fails; }
    }
    std::printf("energy %.15e core %.15e iterations %d nao %d eps0 %.15e epsN %.15e\n", es.GetElectronicEnergy(0),
                es.GetCoreRepulsionEnergy(), es.GetNumberIterationsSCF(), n, eps[0], eps[n - 1]);
    {
        double const* dt = es.GetTotalDipoleMoment();
        std::printf("dipole %.15e %.15e %.15e\n", dt[0], dt[1], dt[2]);
    }
    const double e0 = es.GetElectronicEnergy(0);
    es.DoSCF(false);   // from the converged density: few iterations, same energy
    std::printf("warm_iterations %d warm_energy %.15e\n", es.GetNumberIterationsSCF(), es.GetElectronicEnergy(0));
    {   // forces of the last SCF, one line of 3 N values
        double const* const* F = es.GetForce(0);
        std::printf("forces");
        for (int a = 0; a < natoms; ++a)
            for (int c = 0; c < 3; ++c) std::printf(" %.17g", F[a][c]);
        std::printf("\n");
    }
    if (std::fabs(es.GetElectronicEnergy(0) - e0) > 1e-9) { std::printf("FAIL warm energy\n"); ++fails; }

    nddo::Molecule odd = mol;
    odd.Z.pop_back();
    odd.xyz_bohr.resize(odd.xyz_bohr.size() - 3);
    fails += expect_throw("odd electrons", "Total number of valence electrons is odd. totalNumberValenceElectrons=",
                          [](nddo::NddoElectronicStructure& e, nddo::Molecule& m) { e.SetMolecule(&m); }, es, odd);
    nddo::Molecule bad = mol;
    bad.Z[0] = 16;
    fails += expect_throw("atom type", "Non available atom is contained.",
                          [](nddo::NddoElectronicStructure& e, nddo::Molecule& m) { e.SetMolecule(&m); }, es, bad);
    nddo::NddoElectronicStructure nc((nddo::TheoryType)theory, engine);
    nc.Parameters().SetMaxIterationsSCF(2);
    nc.Parameters().SetThresholdSCF(1e-12);
    fails += expect_throw("not converged", "SCF did not met convergence criterion. maxIterationsSCF=2",
                          [](nddo::NddoElectronicStructure& e, nddo::Molecule& m) {
                              e.SetMolecule(&m);
                              e.DoSCF();
                          },
                          nc, mol);
    std::printf("%s\n", fails ? "FAILED" : "OK");
    return fails ? 1 : 0;
}
