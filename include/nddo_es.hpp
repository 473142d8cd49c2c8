// nddo_es.hpp -- C++ host side of the drop-in: MolDS's ElectronicStructure SCF
// API (src/base/ElectronicStructure.h:23-53) over the C ABI of nddo.h.
//
// Same method names, argument meaning and error behaviour as the reference:
//   SetMolecule(Molecule*)          ElectronicStructure.h:27  (checks as Cndo2::SetMolecule,
//                                   src/cndo/Cndo2.cpp:310-376: odd electron count, atom types)
//   DoSCF(bool requiresGuess=true)  ElectronicStructure.h:28  (Cndo2::DoSCF, Cndo2.cpp:585-729;
//                                   requiresGuess=false starts from the previous density)
//   GetFockMatrix()                 ElectronicStructure.h:30  (after DoSCF its rows are the MO
//                                   coefficient vectors, Lapack.cpp:117-150)
//   GetEnergiesMO()                 ElectronicStructure.h:31
//   GetElectronicEnergy(elecState)  ElectronicStructure.h:38  (ground state only; includes core
//                                   repulsion as Cndo2.cpp:1412-1415)
//   GetCoreRepulsionEnergy()        ElectronicStructure.h:39
//   GetTheoryType()                 ElectronicStructure.h:26
// Errors throw MolDSException (a std::domain_error, src/base/MolDSException.h:26)
// with the reference's messages (Cndo2.cpp:148-157, Lapack.cpp:66).
//
// Header-only; link with -lnddo.  One Engine per HIP device can be shared by
// many NddoElectronicStructure objects (each DoSCF is one nddo_scf_once).
#ifndef MOLDS_AMD_NDDO_ES_HPP
#define MOLDS_AMD_NDDO_ES_HPP

#include <array>
#include <algorithm>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nddo.h"

namespace nddo {

class MolDSException : public std::domain_error {
public:
    MolDSException(const std::string& cause, int status) : std::domain_error(cause), status_(status) {}
    int Status() const { return status_; }   // the nddo.h status code behind the exception

private:
    int status_;
};

enum class TheoryType { MNDO = NDDO_MNDO, AM1 = NDDO_AM1, PM3 = NDDO_PM3 };

// The geometry DoSCF reads: atomic numbers and Cartesian coordinates in bohr
// (MolDS keeps them in Molecule / Atom, src/base/Molecule.h).
struct Molecule {
    std::vector<int> Z;
    std::vector<double> xyz_bohr;   // x, y, z interleaved
    void AddAtom(int z, double x, double y, double zc) {
        Z.push_back(z);
        xyz_bohr.insert(xyz_bohr.end(), {x, y, zc});
    }
    int GetAtomVectSize() const { return (int)Z.size(); }
};

// MolDS's SCF parameters (Parameters::Get*SCF, defaults src/base/Parameters.cpp:132-144).
struct ScfParameters {
    nddo_scf_opts o;
    ScfParameters() { nddo_scf_opts_default(&o); }
    int GetMaxIterationsSCF() const { return o.max_iter; }
    void SetMaxIterationsSCF(int v) { o.max_iter = v; }
    double GetThresholdSCF() const { return o.rms_density; }
    void SetThresholdSCF(double v) { o.rms_density = v; }
    double GetDampingThreshSCF() const { return o.damping_thresh; }
    void SetDampingThreshSCF(double v) { o.damping_thresh = v; }
    double GetDampingWeightSCF() const { return o.damping_weight; }
    void SetDampingWeightSCF(double v) { o.damping_weight = v; }
    int GetDiisNumErrorVectSCF() const { return o.diis_num_error_vect; }
    void SetDiisNumErrorVectSCF(int v) { o.diis_num_error_vect = v; }
    double GetDiisStartErrorSCF() const { return o.diis_start_error; }
    void SetDiisStartErrorSCF(double v) { o.diis_start_error = v; }
    double GetDiisEndErrorSCF() const { return o.diis_end_error; }
    void SetDiisEndErrorSCF(double v) { o.diis_end_error = v; }
};

// RAII owner of one nddo_engine (one HIP device).
class Engine {
public:
    explicit Engine(int hip_device = 0, const nddo_engine_opts* opts = nullptr) {
        const int rc = nddo_engine_create(hip_device, opts, &e_);
        if (rc != NDDO_OK) throw MolDSException(std::string("nddo engine: ") + nddo_last_error(), rc);
    }
    ~Engine() { nddo_engine_destroy(e_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    nddo_engine* get() const { return e_; }

private:
    nddo_engine* e_ = nullptr;
};

class NddoElectronicStructure {
public:
    NddoElectronicStructure(TheoryType theory, std::shared_ptr<Engine> engine, int precision = NDDO_FP64)
        : theory_(theory), engine_(std::move(engine)), precision_(precision) {}

    TheoryType GetTheoryType() const { return theory_; }
    ScfParameters& Parameters() { return params_; }

    void SetMolecule(Molecule* molecule) {
        if (!molecule) throw MolDSException("Error in cndo::Cndo2::DoSCF: A molecule is not set.\n", NDDO_BAD_ARG);
        int n = 0, ne = 0;
        for (int z : molecule->Z) {
            if (!(z == 1 || z == 6 || z == 7 || z == 8 || z == 9))
                throw MolDSException("Error in cndo::Cndo2::CheckEnableAtomType: Non available atom is contained.\n",
                                     NDDO_UNSUPPORTED_ELEMENT);
            n += z == 1 ? 1 : 4;
            ne += z == 1 ? 1 : z - 2;
        }
        if (ne % 2 != 0) {
            std::ostringstream ss;
            ss << "Error in cndo::Cndo2::SetMolecule: Total number of valence electrons is odd. "
                  "totalNumberValenceElectrons="
               << ne << "\n";
            throw MolDSException(ss.str(), NDDO_ODD_ELECTRONS);
        }
        molecule_ = molecule;
        n_ = n;
        fock_.assign((size_t)n * n, 0.0);
        P_.assign((size_t)n * n, 0.0);
        eps_.assign(n, 0.0);
        rows_.resize(n);
        prow_.resize(n);
        for (int i = 0; i < n; ++i) {
            rows_[i] = &fock_[(size_t)i * n];
            prow_[i] = &P_[(size_t)i * n];
        }
        have_density_ = false;
    }

    void DoSCF(bool requiresGuess = true) {
        if (!molecule_) throw MolDSException("Error in cndo::Cndo2::DoSCF: A molecule is not set.\n", NDDO_BAD_ARG);
        nddo_job_desc d{};
        d.theory = (int)theory_;
        d.precision = precision_;
        d.natoms = molecule_->GetAtomVectSize();
        d.Z = molecule_->Z.data();
        d.xyz_bohr = molecule_->xyz_bohr.data();
        d.opts = params_.o;
        d.requires_guess = requiresGuess ? 1 : 0;
        d.P0 = (!requiresGuess && have_density_) ? P_.data() : nullptr;
        std::vector<double> P0;
        if (d.P0) {
            P0 = P_;   // the result overwrites P_
            d.P0 = P0.data();
        }
        nddo_result r{};
        r.eps = eps_.data();
        r.C = fock_.data();
        r.P = P_.data();
        r.dipole = dipole_.data();
        const int st = nddo_scf_once(engine_->get(), &d, &r);
        if (st == NDDO_NOT_CONVERGED) {
            std::ostringstream ss;
            ss << "Error in cndo::Cndo2::DoSCF: SCF did not met convergence criterion. maxIterationsSCF="
               << params_.o.max_iter << "\n";
            throw MolDSException(ss.str(), st);
        }
        if (st == NDDO_EIGENSOLVER_FAILED)
            throw MolDSException("Error in wrappers::Lapack::Dsyevd: info != 0: info = -1\n", st);
        if (st != NDDO_OK) throw MolDSException(std::string("Error in cndo::Cndo2::DoSCF: ") + nddo_last_error(), st);
        have_density_ = true;
        energy_ = r.energy_total;
        e_core_ = r.energy_core;
        hof_ = r.heat_of_formation;
        iterations_ = r.iterations;
    }

    double const* const* GetFockMatrix() const { return rows_.data(); }
    double const* GetEnergiesMO() const { return eps_.data(); }
    double const* const* GetOrbitalElectronPopulation() const { return prow_.data(); }
    double GetElectronicEnergy(int elecState = 0) const {
        if (elecState != 0) throw MolDSException("excited states are not part of this engine", NDDO_BAD_ARG);
        return energy_;
    }
    double GetCoreRepulsionEnergy() const { return e_core_; }
    // Mndo.cpp:395-404 for MNDO; AM1 / PM3 with E_isol from the one-centre parameters
    double GetHeatsFormation() const { return hof_; }
    // ground-state dipole moments (a.u.) as Cndo2::OutputSCFDipole prints them
    // (Cndo2.cpp:748-753, 1126-1184): total, electronic, core
    double const* GetTotalDipoleMoment() const { return dipole_.data(); }
    double const* GetElectronicDipoleMoment() const { return dipole_.data() + 3; }
    double const* GetCoreDipoleMoment() const { return dipole_.data() + 6; }

    // GetForce(elecState) (ElectronicStructure.h:36; Mndo::CalcForce, Mndo.cpp:2711-2986):
    // force[a][c] in Hartree/bohr of the ground state at the last DoSCF's
    // geometry, analytic (nddo_result.forces).  The SCF is first taken from
    // the last DoSCF's density (DoSCF(false)) to density RMS 1e-10, so the
    // forces carry no error from a looser SCF threshold; the adapter's energy,
    // orbitals and density stay those of the last DoSCF.
    double const* const* GetForce(int elecState) {
        if (elecState != 0) throw MolDSException("excited states are not part of this engine", NDDO_BAD_ARG);
        if (!molecule_ || !have_density_)
            throw MolDSException("Error in cndo::Cndo2::GetForce: DoSCF has not run.\n", NDDO_BAD_ARG);
        const int N = molecule_->GetAtomVectSize();
        nddo_job_desc d{};
        d.theory = (int)theory_;
        d.precision = precision_;
        d.natoms = N;
        d.Z = molecule_->Z.data();
        d.xyz_bohr = molecule_->xyz_bohr.data();
        d.opts = params_.o;
        d.opts.rms_density = std::min(d.opts.rms_density, 1.0e-10);
        d.opts.max_iter = std::max(d.opts.max_iter, 200);
        d.requires_guess = 0;
        std::vector<double> P0 = P_, eps(n_);
        d.P0 = P0.data();
        force_.assign((size_t)3 * N, 0.0);
        nddo_result r{};
        r.eps = eps.data();
        r.forces = force_.data();
        const int st = nddo_scf_once(engine_->get(), &d, &r);
        if (st != NDDO_OK)
            throw MolDSException(std::string("Error in GetForce: the SCF for the forces ended with: ") + nddo_last_error(),
                                 st);
        frow_.resize(N);
        for (int a = 0; a < N; ++a) frow_[a] = &force_[(size_t)3 * a];
        return frow_.data();
    }
    int GetNumberIterationsSCF() const { return iterations_; }
    int GetTotalNumberAOs() const { return n_; }

private:
    TheoryType theory_;
    std::shared_ptr<Engine> engine_;
    int precision_;
    ScfParameters params_;
    Molecule* molecule_ = nullptr;
    int n_ = 0;
    std::vector<double> fock_, P_, eps_;
    std::vector<double*> rows_, prow_;
    std::vector<double> force_;
    std::vector<double*> frow_;
    bool have_density_ = false;
    double energy_ = 0.0, e_core_ = 0.0, hof_ = 0.0;
    std::array<double, 9> dipole_{};
    int iterations_ = 0;
};

}  // namespace nddo

#endif  // MOLDS_AMD_NDDO_ES_HPP
