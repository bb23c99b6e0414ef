// C++ facade over nnp_capi.h for in-tree C++ callers, with the semantics of AtomNN's potential
// plug-in: ptnl::Pot (src/torch/pot.hpp:29-107) as PotNNPE implements it
// (src/torch/pot_nnpe.hpp:22-64, pot_nnpe.cpp:49-300):
//   * energy(...)  <- PotNNPE::energy(const Structure&, const NeighborList&): the total energy,
//                     the species' type energies included (pot_nnpe.cpp:49-106);
//   * compute(...) <- PotNNPE::compute(Structure&, const NeighborList&): returns that energy and
//                     ADDS the forces into the caller's array (struc.force(i) +=, pot_nnpe.cpp:
//                     175-300); the reaction of a pair goes to its minimum image only
//                     (pot_nnpe.cpp:250-285), or to every image with all_images;
//   * index(name)  <- PotNNPE::index(const char*) (pot_nnpe.hpp:37), -1 when absent.
// The neighbour list is built on the GPU (NeighborList::build, src/struc/neighbor.cpp:59-171),
// so no list is passed.  Errors become std::runtime_error with the message of nnp_last_error().
// Header only; link with -latomnn_amd.  One host thread per object at a time.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "nnp_capi.h"

namespace atomnn {

class GpuPot {
  public:
    explicit GpuPot(const std::string& ann_path) {
        if (nnp_load(ann_path.c_str(), &m_) != NNP_OK) throw std::runtime_error(nnp_last_error());
    }
    ~GpuPot() { nnp_free(m_); }
    GpuPot(const GpuPot&) = delete;
    GpuPot& operator=(const GpuPot&) = delete;

    int ntypes() const { return nnp_ntypes(m_); }
    double rc() const { return nnp_rc(m_); }
    int index(const std::string& name) const { return nnp_type_index(m_, name.c_str()); }
    nnp_model* handle() { return m_; }

    // Total energy (type energies included) of n atoms: type[n] species indices, pos[3n]
    // Cartesian positions (A), cell[9] lattice vectors a, b, c or nullptr (open).
    double energy(int n, const int32_t* type, const double* pos, const double* cell) {
        return run(n, type, pos, cell, nullptr, false);
    }
    // The same energy; force[3n] += the forces (eV/A).
    double compute(int n, const int32_t* type, const double* pos, const double* cell, double* force,
                   bool all_images = false) {
        return run(n, type, pos, cell, force, all_images);
    }
    // Many structures in one packed pass (the per-structure loop of nnpce.cpp:273-281): energies
    // into energy[k], per-atom energies into eatom[k] (may be empty) and forces into force[k]
    // (overwritten; may be empty for energy only).
    void compute_batch(const std::vector<nnp_structure>& s, std::vector<double>& energy,
                       std::vector<double*>& eatom, std::vector<double*>& force) {
        const size_t ns = s.size();
        if (ns == 0) return;
        std::vector<nnp_result> r(ns);
        const bool want_f = !force.empty(), want_e = !eatom.empty();
        for (size_t k = 0; k < ns; ++k) r[k] = nnp_result{0.0, want_e ? eatom[k] : nullptr, want_f ? force[k] : nullptr};
        const uint32_t flags = NNP_VACUUM_ENERGY | (want_e ? NNP_ATOM_ENERGY : 0u) | (want_f ? NNP_FORCE : 0u);
        if (nnp_compute_batch(m_, (int)ns, s.data(), flags, r.data()) != NNP_OK) throw std::runtime_error(nnp_last_error());
        energy.resize(ns);
        for (size_t k = 0; k < ns; ++k) energy[k] = r[k].energy;
    }

  private:
    double run(int n, const int32_t* type, const double* pos, const double* cell, double* force, bool all_images) {
        nnp_structure s{n, type, pos, cell};
        if (force) f_.assign(3 * (size_t)(n > 0 ? n : 0), 0.0);
        nnp_result r{0.0, nullptr, force ? f_.data() : nullptr};
        const uint32_t flags =
            NNP_VACUUM_ENERGY | (force ? NNP_FORCE : 0u) | (all_images ? NNP_ALL_IMAGES : 0u);
        if (nnp_compute(m_, &s, flags, &r) != NNP_OK) throw std::runtime_error(nnp_last_error());
        if (force)
            for (size_t q = 0; q < f_.size(); ++q) force[q] += f_[q];
        return r.energy;
    }

    nnp_model* m_ = nullptr;
    std::vector<double> f_;
};

}  // namespace atomnn
