// H_psi_hip.cppm — C++23 module interface unit over the facade
// (effm_hamiltonian.hpp), so that reference callers keep `import`:
//
//   import H_psi.hip;          // instead of: import H_psi;
//   Hamiltonian h(cell);       // the reference's class and method names
//   h.at_k(ikpt)(psi, hpsi);   // hpsi += H psi
//
// Reference interface: src/H_psi/hamiltonian.cppm:115-363 (module
// H_psi.hamiltonian, re-exported by src/H_psi.cppm).  The names are exported
// both in namespace effm and, as the reference spells them, at global scope.
// Build: clang++ -std=c++23 --precompile H_psi_hip.cppm (libstdc++ headers in
// the global module fragment), then link libeffm_hip.so.
module;
#include "effm_hamiltonian.hpp"
export module H_psi.hip;

export namespace effm {
using effm::CellData;
using effm::GradHPsi;
using effm::GradHPsi32;
using effm::Hamiltonian;
using effm::HessHPsi;
using effm::KPointData;
using effm::Precision;
using effm::SpeciesData;
using effm::throw_status;
}  // namespace effm

export using effm::GradHPsi;
export using effm::Hamiltonian;
export using effm::HessHPsi;
